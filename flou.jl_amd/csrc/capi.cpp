// C ABI of libflouhip.so (declared in include/flouhip.h): validation of the flattened
// MultielementDisc, upload of operators / connectivity / geometry, per-element face
// tables, kernel dispatch and the fixed-dt time loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flouhip.h"
#include "common.hpp"
#include "fast.hpp"
#include "reduce.hpp"

using namespace flouhip;

struct flou_hip_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int nd = 0, n = 0, nv = 0, prec = 0, npe = 0, npf = 0, nbd = 0;
    int64_t ne = 0, nf = 0, ndofs = 0;
    int eq = 0, vol = 0, tp = 0, savg = 0, sdiss = 0, typo = 0;
    double gamma = 1.4, intensity = 1.0, a[3] = {0, 0, 0}, anorm = 0.0;
    int gather = 0;
    int lift_end = 0;
    int egeom_stride = 0;
    int curved = 0;   // per-node geometry (GEOM_CURVED)
    double blend = 1.0, sjac[3] = {1.0, 1.0, 1.0};  // HybridDivOperator
    int int_npe = 1;  // node stride of the integrate weights (1: per-node Jω)
    int lim_kind = FLOU_HIP_LIMITER_NONE;  // stage limiter applied by flou_hip_rk_stage
    double lim_min = 0.0;
    double* d_evol = nullptr;  // element volumes (limiter means)
    double* d_eqA = nullptr;   // equispaced interpolation matrix of the last output call
    std::vector<double> eqA;   // its host copy (source of the pending upload)
    // device tables
    void* d_ngeom = nullptr;
    void* d_ops = nullptr;
    int* d_nb = nullptr;
    int* d_finfo = nullptr;
    int* d_fgi = nullptr;
    int* d_fid = nullptr;
    void* d_egeom = nullptr;
    void* d_fgeom = nullptr;
    double* d_dxe = nullptr;
    double* d_jac = nullptr;
    double* d_w = nullptr;
    void* d_bcstate = nullptr;
    void* d_bcconst = nullptr;
    int* d_bctype = nullptr;
    int* d_err = nullptr;
    double* d_partial = nullptr;
    double* d_out = nullptr;
    double* h_out = nullptr;  // pinned
    // host mirrors of the boundary data
    std::vector<int> bctype;
    std::vector<double> bcstate, bcconst;
    // time-loop work buffers
    void* d_work[3] = {nullptr, nullptr, nullptr};
    // this device's address of the static scratch array (flou_hip_rk_step_vjp_noalloc)
    void* d_scratch = nullptr;
    std::string scratch_why;
    // kernels
    RhsLauncherF64 l64 = nullptr;
    RhsLauncherF32 l32 = nullptr;
    VjpLauncherF64 v64 = nullptr;
    VjpLauncherF32 v32 = nullptr;
    std::string vjp_why;  // why the adjoint is unavailable for this handle
    // FP32 fast path (2-D/3-D split form on axis-aligned affine or curved elements)
    FastLauncher lfast = nullptr;
    int force_plain = 0;
    void* d_fegeom = nullptr;
    int fegeom_stride = 0;
    void* d_fdsharp = nullptr;
    void* d_fnj = nullptr;      // curved fast kernel: per face node (normal, surface J), float4
    // FP32 adjoint on the fast path: its kernel, S_ij = ω_i D♯_ij and 1/ω
    FastVjpLauncher vfast = nullptr;
    void* d_fwsh = nullptr;
    void* d_fiw = nullptr;
    float lift_m = 0.0f, lift_p = 0.0f, fdiss = 0.0f;
    std::string err;
};

// Work states of flou_hip_rk_step_vjp_noalloc: three state buffers of one reverse RK step
// (FLOU_HIP_NOALLOC_SCRATCH_BYTES).  Each GPU that the library runs on holds its own copy
// until the process ends.  In host code `scratch` only names the symbol: kernels receive
// the device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[FLOU_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` on `device` (which must be current), looked up once per device.
hipError_t scratch_address(int device, void** out) {
    static std::mutex mu;
    static std::vector<void*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= addr.size()) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = p;
    }
    *out = addr[device];
    return hipSuccess;
}

int fail(flou_hip_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHECK(h, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(h, e_ == hipErrorOutOfMemory ? FLOU_HIP_ENOMEM : FLOU_HIP_EDEVICE,  \
                        std::string(#call) + ": " + hipGetErrorString(e_));                 \
    } while (0)

template <typename T>
int upload(flou_hip_handle* h, void** dptr, const std::vector<T>& v) {
    if (v.empty()) return 0;
    HIPCHECK(h, hipMalloc(dptr, v.size() * sizeof(T)));
    HIPCHECK(h, hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// Upload a double vector converted to the handle's precision.
int upload_real(flou_hip_handle* h, void** dptr, const std::vector<double>& v) {
    if (h->prec == FLOU_HIP_FP64) return upload(h, dptr, v);
    std::vector<float> f(v.begin(), v.end());
    return upload(h, dptr, f);
}

void free_all(flou_hip_handle* h) {
    void* ptrs[] = {h->d_ops, h->d_nb, h->d_finfo, h->d_fgi, h->d_fid, h->d_egeom, h->d_fgeom,
                    h->d_dxe, h->d_jac, h->d_w, h->d_bcstate, h->d_bcconst, h->d_bctype,
                    h->d_err, h->d_partial, h->d_out, h->d_work[0], h->d_work[1], h->d_work[2],
                    h->d_fegeom, h->d_fdsharp, h->d_ngeom, h->d_evol, h->d_eqA, h->d_fnj,
                    h->d_fwsh, h->d_fiw};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->h_out) (void)hipHostFree(h->h_out);
}

int bytes_of(const flou_hip_handle* h) { return h->prec == FLOU_HIP_FP64 ? 8 : 4; }

// 2N tableaux (Carpenter & Kennedy, NASA TM-109112 (1994); Allampalli et al., JCP 228 (2009)
// as tabulated by OrdinaryDiffEq 6.49.1).
bool tableau(int scheme, int stage, double& A, double& B) {
    static const double ckA[5] = {0.0, -567301805773.0 / 1357537059087.0,
                                  -2404267990393.0 / 2016746695238.0,
                                  -3550918686646.0 / 2091501179385.0,
                                  -1275806237668.0 / 842570457699.0};
    static const double ckB[5] = {1432997174477.0 / 9575080441755.0,
                                  5161836677717.0 / 13612068292357.0,
                                  1720146321549.0 / 2090206949498.0,
                                  3134564353537.0 / 4481467310338.0,
                                  2277821191437.0 / 14882151754819.0};
    static const double orkA[5] = {0.0, -1.0, -1.55798, -1.0, -0.45031};
    static const double orkB[5] = {0.2, 0.83204, 0.6, 0.35394, 0.2};
    if (stage < 0 || stage > 4) return false;
    if (scheme == FLOU_HIP_CK2N54) { A = ckA[stage]; B = ckB[stage]; return true; }
    if (scheme == FLOU_HIP_ORK256) { A = orkA[stage]; B = orkB[stage]; return true; }
    return false;
}

std::string& create_error() {
    static thread_local std::string msg;
    return msg;
}

template <typename T>
RhsArgs<T> make_args(flou_hip_handle* h) {
    RhsArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.ndofs = h->ndofs;
    a.ne = (int)h->ne;
    a.nb = h->d_nb;
    a.finfo = h->d_finfo;
    a.fgi = h->d_fgi;
    a.fid = h->d_fid;
    a.egeom = (const T*)h->d_egeom;
    a.egeom_stride = h->egeom_stride;
    a.fgeom = (const T*)h->d_fgeom;
    a.ngeom = (const T*)h->d_ngeom;
    a.ops = (const T*)h->d_ops;
    a.bcstate = (const T*)h->d_bcstate;
    a.bcconst = (const T*)h->d_bcconst;
    a.bctype = h->d_bctype;
    a.gather = h->gather;
    a.lift_end = h->lift_end;
    a.gamma = (T)h->gamma;
    a.intensity = (T)h->intensity;
    a.a0 = (T)h->a[0];
    a.a1 = (T)h->a[1];
    a.a2 = (T)h->a[2];
    a.savg = h->savg;
    a.sdiss = h->sdiss;
    a.typo = h->typo;
    a.errflag = h->d_err;
    a.blend = (T)h->blend;
    for (int k = 0; k < 3; ++k) a.sjac[k] = (T)h->sjac[k];
    return a;
}

// The fast kernels' arguments for the state Q (the stage fields left zero).
FastArgs fast_args(const flou_hip_handle* h, const void* Q) {
    FastArgs a;
    std::memset(&a, 0, sizeof(a));
    a.Q = (const float*)Q;
    a.ndofs = h->ndofs;
    a.ne = (int)h->ne;
    a.nb = h->d_nb;
    a.finfo = h->d_finfo;
    a.fgi = h->d_fgi;
    a.fid = h->d_fid;
    a.egeom = (const float*)h->d_fegeom;
    a.egeom_stride = h->fegeom_stride;
    a.fgeom = (const float*)h->d_fgeom;
    a.fnj = (const float4*)h->d_fnj;
    a.ngeom = (const float*)h->d_ngeom;
    a.dsharp = (const float*)h->d_fdsharp;
    a.lift_m = h->lift_m;
    a.lift_p = h->lift_p;
    a.bcstate = (const float*)h->d_bcstate;
    a.bcconst = (const float*)h->d_bcconst;
    a.bctype = h->d_bctype;
    a.gamma = (float)h->gamma;
    a.inv_gm1 = (float)(1.0 / (h->gamma - 1.0));
    a.diss = h->fdiss;
    a.errflag = h->d_err;
    return a;
}

// One adjoint pass (see VjpArgs): o0 = a0·P0 + a1·P1 + (d rhs/dQ)^T(c·Y) at Q,
// o1 = b0·o0 + b1·P1 + b2·P2.  o0/o1 may alias P0/P1/P2 (point-wise) but not Y or Q.
int launch_vjp(flou_hip_handle* h, const void* Q, const void* Y, double c, const void* P0, double a0,
               const void* P1, double a1, void* o0, const void* P2, double b0, double b1, double b2,
               void* o1) {
    if (!h->v64 && !h->v32) return fail(h, FLOU_HIP_EUNSUPPORTED, h->vjp_why);
    hipError_t e;
    if (h->prec == FLOU_HIP_FP64) {
        VjpArgs<double> a;
        std::memset(&a, 0, sizeof(a));
        a.r = make_args<double>(h);
        a.r.Q = (const double*)Q;
        a.Y = (const double*)Y; a.c = c;
        a.P0 = (const double*)P0; a.P1 = (const double*)P1; a.P2 = (const double*)P2;
        a.a0 = a0; a.a1 = a1; a.b0 = b0; a.b1 = b1; a.b2 = b2;
        a.o0 = (double*)o0; a.o1 = (double*)o1;
        e = h->v64(a, h->stream);
    } else if (h->vfast && !h->force_plain) {
        FastVjpArgs a;
        std::memset(&a, 0, sizeof(a));
        a.f = fast_args(h, Q);
        a.Y = (const float*)Y; a.c = (float)c;
        a.P0 = (const float*)P0; a.P1 = (const float*)P1; a.P2 = (const float*)P2;
        a.a0 = (float)a0; a.a1 = (float)a1; a.b0 = (float)b0; a.b1 = (float)b1; a.b2 = (float)b2;
        a.o0 = (float*)o0; a.o1 = (float*)o1;
        a.wsh = (const float*)h->d_fwsh;
        a.iw = (const float*)h->d_fiw;
        e = h->vfast(a, h->stream);
    } else {
        VjpArgs<float> a;
        std::memset(&a, 0, sizeof(a));
        a.r = make_args<float>(h);
        a.r.Q = (const float*)Q;
        a.Y = (const float*)Y; a.c = (float)c;
        a.P0 = (const float*)P0; a.P1 = (const float*)P1; a.P2 = (const float*)P2;
        a.a0 = (float)a0; a.a1 = (float)a1; a.b0 = (float)b0; a.b1 = (float)b1; a.b2 = (float)b2;
        a.o0 = (float*)o0; a.o1 = (float*)o1;
        e = h->v32(a, h->stream);
    }
    if (e != hipSuccess) return fail(h, FLOU_HIP_EDEVICE, std::string("adjoint kernel launch: ") + hipGetErrorString(e));
    return FLOU_HIP_OK;
}

int launch_stage(flou_hip_handle* h, int mode, int first, double dt, double A, double B,
                 const void* u_in, const void* ubase, void* reg, void* out) {
    hipError_t e;
    if (h->prec == FLOU_HIP_FP64) {
        RhsArgs<double> a = make_args<double>(h);
        a.Q = (const double*)u_in;
        a.out = (double*)out;
        a.ubase = (const double*)ubase;
        a.reg = (double*)reg;
        a.dt = dt; a.A = A; a.B = B;
        a.mode = mode; a.first = first;
        e = h->l64(a, h->stream);
    } else if (h->lfast && !h->force_plain) {
        FastArgs a = fast_args(h, u_in);
        a.out = (float*)out;
        a.ubase = (const float*)ubase;
        a.reg = (float*)reg;
        a.dt = (float)dt; a.A = (float)A; a.B = (float)B;
        a.mode = mode; a.first = first;
        e = h->lfast(a, h->stream);
    } else {
        RhsArgs<float> a = make_args<float>(h);
        a.Q = (const float*)u_in;
        a.out = (float*)out;
        a.ubase = (const float*)ubase;
        a.reg = (float*)reg;
        a.dt = (float)dt; a.A = (float)A; a.B = (float)B;
        a.mode = mode; a.first = first;
        e = h->l32(a, h->stream);
    }
    if (e != hipSuccess) return fail(h, FLOU_HIP_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return FLOU_HIP_OK;
}

bool overlaps(const void* a, const void* b, size_t bytes) {
    const char* x = (const char*)a;
    const char* y = (const char*)b;
    return x < y + bytes && y < x + bytes;
}

}  // namespace

extern "C" {

const char* flou_hip_version(void) { return "flouhip 0.1.0 (gfx950)"; }

const char* flou_hip_last_error(const flou_hip_handle* h) { return h ? h->err.c_str() : "null handle"; }

const char* flou_hip_create_error(void) { return create_error().c_str(); }

int flou_hip_create(const flou_hip_desc* d, flou_hip_handle** out) {
    if (!d || !out) return FLOU_HIP_EINVAL;
    *out = nullptr;
    flou_hip_handle* h = new flou_hip_handle();
    auto bad = [&](int code, const std::string& msg) {
        // the handle is not returned: keep the message for flou_hip_create_error()
        create_error() = msg;
        free_all(h);
        delete h;
        return code;
    };
    create_error().clear();
    if (d->nd < 1 || d->nd > 3) return bad(FLOU_HIP_EINVAL, "nd must be 1, 2 or 3");
    if (d->n < 2 || d->n > 8) return bad(FLOU_HIP_EUNSUPPORTED, "n (nodes per direction) must be in 2..8");
    if (d->precision != FLOU_HIP_FP64 && d->precision != FLOU_HIP_FP32)
        return bad(FLOU_HIP_EINVAL, "unknown precision");
    if (d->equation == FLOU_HIP_EULER) {
        if (d->nvars != d->nd + 2) return bad(FLOU_HIP_EINVAL, "Euler needs nvars = nd + 2");
        if (!(d->gamma > 1.0)) return bad(FLOU_HIP_EINVAL, "gamma must be > 1");
    } else if (d->equation == FLOU_HIP_ADVECTION) {
        if (d->nvars != 1) return bad(FLOU_HIP_EINVAL, "advection needs nvars = 1");
    } else {
        return bad(FLOU_HIP_EINVAL, "unknown equation");
    }
    if (d->volume_op != FLOU_HIP_STRONG && d->volume_op != FLOU_HIP_SPLIT && d->volume_op != FLOU_HIP_HYBRID)
        return bad(FLOU_HIP_EINVAL, "unknown volume operator");
    if (d->volume_op == FLOU_HIP_HYBRID) {
        if (d->equation != FLOU_HIP_EULER)
            return bad(FLOU_HIP_EUNSUPPORTED, "HybridDivOperator: only the Euler equations are supported");
        if (d->geometry != FLOU_HIP_GEOM_AFFINE)
            return bad(FLOU_HIP_EUNSUPPORTED, "HybridDivOperator: subcell geometry exists for Cartesian elements only");
        if (!std::isfinite(d->blend) || d->blend < 0.0)
            return bad(FLOU_HIP_EINVAL, "HybridDivOperator: the blending constant must be finite and >= 0");
    }
    if (d->tp_flux < 0 || d->tp_flux > 2) return bad(FLOU_HIP_EINVAL, "unknown two-point flux");
    if (d->surface_avg < 0 || d->surface_avg > 1 || d->surface_diss < 0 || d->surface_diss > 3)
        return bad(FLOU_HIP_EINVAL, "unknown surface flux");
    if (d->equation == FLOU_HIP_ADVECTION) {
        if (d->volume_op == FLOU_HIP_SPLIT && d->tp_flux != FLOU_HIP_TP_STDAVG)
            return bad(FLOU_HIP_EINVAL, "advection has only the StdAverage two-point flux");
        if (d->surface_avg != FLOU_HIP_AVG_STD || d->surface_diss >= FLOU_HIP_DISS_MATRIX)
            return bad(FLOU_HIP_EINVAL, "advection has only StdAverage and LxF surface fluxes");
    }
    if (d->nelements < 1 || d->nelements > (int64_t)1 << 30) return bad(FLOU_HIP_EINVAL, "bad nelements");
    if (d->nfaces < 1 || d->nfaces > (int64_t)1 << 31) return bad(FLOU_HIP_EINVAL, "bad nfaces");
    if (d->nboundaries < 0) return bad(FLOU_HIP_EINVAL, "bad nboundaries");
    if (d->geometry != FLOU_HIP_GEOM_AFFINE && d->geometry != FLOU_HIP_GEOM_CURVED)
        return bad(FLOU_HIP_EINVAL, "unknown geometry kind");
    if (d->geometry == FLOU_HIP_GEOM_CURVED && d->nd == 1)
        return bad(FLOU_HIP_EINVAL, "curved geometry needs nd >= 2");
    if (!d->D || !d->Ds || !d->Dsharp || !d->l_minus || !d->l_plus || !d->dg_minus || !d->dg_plus || !d->omega)
        return bad(FLOU_HIP_EINVAL, "missing operator arrays");
    if (!d->elem_faceinds || !d->elem_facepos || !d->face_eleminds || !d->face_elempos)
        return bad(FLOU_HIP_EINVAL, "missing connectivity arrays");
    if (!d->elem_jac || !d->elem_metric || !d->elem_volume || !d->face_frames || !d->face_jac)
        return bad(FLOU_HIP_EINVAL, "missing geometry arrays");

    const int nd = d->nd, n = d->n, NF = 2 * nd;
    h->nd = nd; h->n = n; h->nv = d->nvars; h->prec = d->precision;
    h->ne = d->nelements; h->nf = d->nfaces; h->nbd = d->nboundaries;
    h->npe = 1; h->npf = 1;
    for (int k = 0; k < nd; ++k) h->npe *= n;
    for (int k = 0; k < nd - 1; ++k) h->npf *= n;
    h->ndofs = h->ne * h->npe;
    h->eq = d->equation; h->vol = d->volume_op; h->tp = d->tp_flux;
    h->savg = d->surface_avg; h->sdiss = d->surface_diss;
    h->gamma = d->gamma; h->intensity = d->intensity;
    h->blend = d->volume_op == FLOU_HIP_HYBRID ? d->blend : 1.0;
    for (int k = 0; k < 3; ++k) h->a[k] = d->equation == FLOU_HIP_ADVECTION ? d->velocity[k] : 0.0;
    {
        double a2 = h->a[0] * h->a[0];
        for (int k = 1; k < nd; ++k) a2 = a2 + h->a[k] * h->a[k];
        h->anorm = std::sqrt(a2);
    }
    h->typo = 1;  // as the reference computes (option 0)

    // connectivity checks and per element-face tables
    const int64_t ne = h->ne, nf = h->nf;
    // face orientations (master2slave, StdSegment.jl:161-167, StdQuad.jl:161-181): 0 or 1 on
    // segment faces, 0..7 on quad faces, 0 on point faces and on boundary faces
    bool oriented = false;
    if (d->face_orientation)
        for (int64_t f = 0; f < nf; ++f) {
            const int o = d->face_orientation[f];
            if (o >= (nd == 3 ? 8 : (nd == 2 ? 2 : 1)))
                return bad(FLOU_HIP_EINVAL, "face orientation out of range for the face's shape");
            if (o != 0 && d->face_eleminds[f * 2 + 1] < 0)
                return bad(FLOU_HIP_EINVAL, "boundary faces must have orientation 0");
            oriented |= o != 0;
        }
    std::vector<int> nb(ne * NF), finfo(ne * NF), fgi(ne * NF), fid(ne * NF);
    for (int64_t e = 0; e < ne; ++e)
        for (int k = 0; k < NF; ++k) {
            const int64_t f = d->elem_faceinds[e * NF + k];
            const int s = d->elem_facepos[e * NF + k];
            if (f < 0 || f >= nf || (s != 0 && s != 1)) return bad(FLOU_HIP_EINVAL, "bad elem_faceinds/elem_facepos");
            if (d->face_eleminds[f * 2 + s] != e || d->face_elempos[f * 2 + s] != k)
                return bad(FLOU_HIP_EINVAL, "element and face connectivity disagree");
            const int64_t other = d->face_eleminds[f * 2 + 1 - s];
            int info = s;
            if (other >= 0) {
                if (other >= ne) return bad(FLOU_HIP_EINVAL, "bad face_eleminds");
                const int kp = d->face_elempos[f * 2 + 1 - s];
                if (kp < 0 || kp >= NF) return bad(FLOU_HIP_EINVAL, "bad face_elempos");
                info |= kp << 1;
                if (d->face_orientation) info |= (int)d->face_orientation[f] << 28;
            } else {
                if (s != 0) return bad(FLOU_HIP_EINVAL, "boundary faces must have the element on side 1");
                const int bd = d->face_boundary ? d->face_boundary[f] : -1;
                if (bd < 0 || bd >= h->nbd) return bad(FLOU_HIP_EINVAL, "boundary face without a valid boundary index");
                if (bd >= (1 << 24) - 1) return bad(FLOU_HIP_EINVAL, "too many boundaries");
                info |= (bd + 1) << 4;
            }
            nb[e * NF + k] = (int)other;
            finfo[e * NF + k] = info;
            fid[e * NF + k] = (int)f;
        }

    // operators
    std::vector<double> ops(ops_size(n));
    std::memcpy(&ops[ops_D(n)], d->D, n * n * sizeof(double));
    std::memcpy(&ops[ops_Ds(n)], d->Ds, n * n * sizeof(double));
    std::memcpy(&ops[ops_Dsh(n)], d->Dsharp, n * n * sizeof(double));
    std::memcpy(&ops[ops_lm(n)], d->l_minus, n * sizeof(double));
    std::memcpy(&ops[ops_lp(n)], d->l_plus, n * sizeof(double));
    std::memcpy(&ops[ops_gm(n)], d->dg_minus, n * sizeof(double));
    std::memcpy(&ops[ops_gp(n)], d->dg_plus, n * sizeof(double));
    std::memcpy(&ops[ops_w(n)], d->omega, n * sizeof(double));
    h->gather = 1;
    h->lift_end = 1;
    for (int j = 0; j < n; ++j) {
        if (d->l_minus[j] != (j == 0 ? 1.0 : 0.0)) h->gather = 0;
        if (d->l_plus[j] != (j == n - 1 ? 1.0 : 0.0)) h->gather = 0;
        if ((j > 0 && d->dg_minus[j] != 0.0) || (j < n - 1 && d->dg_plus[j] != 0.0)) h->lift_end = 0;
    }

    // geometry
    const int NG = nd * nd + 1;
    const int FG = 3 * nd + 1;
    std::vector<double> eg, fg, ng;
    bool uniform = false;  // affine: every element has the same metric and J
    h->curved = d->geometry == FLOU_HIP_GEOM_CURVED;
    if (h->curved) {
        // per node: metric (column-major nd x nd) and J, stored component-major
        // [NG][ndofs] so that a thread per node reads coalesced; per face node: frame + jac.
        const int64_t N = h->ndofs;
        ng.resize((size_t)NG * N);
        for (int64_t i = 0; i < N; ++i) {
            if (!(d->elem_jac[i] > 0.0)) return bad(FLOU_HIP_EINVAL, "node Jacobians must be positive");
            for (int c = 0; c < nd * nd; ++c) ng[c * N + i] = d->elem_metric[i * nd * nd + c];
            ng[(size_t)nd * nd * N + i] = d->elem_jac[i];
        }
        eg.resize(ne * NG);  // per element: node 0 (only read as a placeholder)
        for (int64_t e = 0; e < ne; ++e)
            for (int c = 0; c < NG; ++c) eg[e * NG + c] = ng[(size_t)c * N + e * h->npe];
        h->egeom_stride = NG;
        fg.resize((size_t)nf * h->npf * FG);
        for (int64_t q = 0; q < nf * h->npf; ++q) {
            for (int c = 0; c < 3 * nd; ++c) fg[q * FG + c] = d->face_frames[q * 3 * nd + c];
            fg[q * FG + 3 * nd] = d->face_jac[q];
        }
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < NF; ++k) fgi[e * NF + k] = (int)fid[e * NF + k];
    } else {
        uniform = true;
        for (int64_t e = 1; e < ne && uniform; ++e) {
            if (d->elem_jac[e] != d->elem_jac[0]) uniform = false;
            for (int c = 0; c < nd * nd; ++c)
                if (d->elem_metric[e * nd * nd + c] != d->elem_metric[c]) uniform = false;
        }
        if (h->vol == FLOU_HIP_HYBRID && !uniform)
            return bad(FLOU_HIP_EUNSUPPORTED, "HybridDivOperator: the elements must share one geometry (CartesianMesh)");
        if (h->vol == FLOU_HIP_HYBRID)  // subcell Jacobians = diagonal metric (PhysicalRegions.jl:75-145)
            for (int k = 0; k < nd; ++k) h->sjac[k] = d->elem_metric[k * nd + k];
        if (uniform) {
            eg.resize(NG);
            for (int c = 0; c < nd * nd; ++c) eg[c] = d->elem_metric[c];
            eg[nd * nd] = d->elem_jac[0];
            h->egeom_stride = 0;
        } else {
            eg.resize(ne * NG);
            for (int64_t e = 0; e < ne; ++e) {
                for (int c = 0; c < nd * nd; ++c) eg[e * NG + c] = d->elem_metric[e * nd * nd + c];
                eg[e * NG + nd * nd] = d->elem_jac[e];
            }
            h->egeom_stride = NG;
        }
        for (int64_t e = 0; e < ne; ++e)
            if (!(d->elem_jac[e] > 0.0)) return bad(FLOU_HIP_EINVAL, "element Jacobians must be positive");
        // faces: frames + jac per master position when every face with that position agrees
        std::vector<int> rep(NF, -1);
        bool funiform = true;
        for (int64_t f = 0; f < nf && funiform; ++f) {
            const int pos = d->face_elempos[f * 2];
            if (pos < 0 || pos >= NF) return bad(FLOU_HIP_EINVAL, "bad master face position");
            if (rep[pos] < 0) { rep[pos] = (int)f; continue; }
            const int64_t r = rep[pos];
            if (d->face_jac[f] != d->face_jac[r]) funiform = false;
            for (int c = 0; c < 3 * nd; ++c)
                if (d->face_frames[f * 3 * nd + c] != d->face_frames[r * 3 * nd + c]) funiform = false;
        }
        if (funiform) {
            fg.assign(NF * FG, 0.0);
            for (int p = 0; p < NF; ++p) {
                if (rep[p] < 0) continue;
                for (int c = 0; c < 3 * nd; ++c) fg[p * FG + c] = d->face_frames[(int64_t)rep[p] * 3 * nd + c];
                fg[p * FG + 3 * nd] = d->face_jac[rep[p]];
            }
        } else {
            fg.resize(nf * FG);
            for (int64_t f = 0; f < nf; ++f) {
                for (int c = 0; c < 3 * nd; ++c) fg[f * FG + c] = d->face_frames[f * 3 * nd + c];
                fg[f * FG + 3 * nd] = d->face_jac[f];
            }
        }
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < NF; ++k) {
                const int64_t f = fid[e * NF + k];
                fgi[e * NF + k] = funiform ? d->face_elempos[f * 2] : (int)f;
            }
    }

    // reductions: Δx per element and quadrature weights (FP64)
    std::vector<double> dxe(ne), jac(h->curved ? 0 : ne), w(h->npe);
    for (int64_t e = 0; e < ne; ++e) {
        double dx = d->elem_volume[e] / h->npe;
        dxe[e] = nd == 1 ? dx : (nd == 2 ? std::sqrt(dx) : std::cbrt(dx));
        if (!h->curved) jac[e] = d->elem_jac[e];
    }
    for (int i = 0; i < h->npe; ++i) {
        int idx = i;
        double wi[3];
        for (int k = 0; k < nd; ++k) { wi[k] = d->omega[idx % n]; idx /= n; }
        double x = wi[0];
        for (int k = 1; k < nd; ++k) x = x * wi[k];
        w[i] = x;
    }
    h->int_npe = h->npe;
    if (h->curved) {  // integrate with per-node weights J_i ω_i (one rounding, as Σ (Jω)_i Q_i)
        jac.resize(h->ndofs);
        for (int64_t i = 0; i < h->ndofs; ++i) jac[i] = d->elem_jac[i] * w[i % h->npe];
        w.assign(1, 1.0);
        h->int_npe = 1;
    }

    // kernels
    // the split and hybrid forms on nodes without boundaries (Gauss) have entropy-projected
    // surface terms
    const int kvol = h->gather ? h->vol
                     : h->vol == FLOU_HIP_SPLIT ? (int)VOL_SPLIT_NB
                     : h->vol == FLOU_HIP_HYBRID ? (int)VOL_HYBRID_NB
                                                 : h->vol;
    if (kvol == VOL_SPLIT_NB && (h->eq != FLOU_HIP_EULER || h->curved))
        return bad(FLOU_HIP_EUNSUPPORTED,
                   "the split form on Gauss nodes is implemented for the Euler equations on Cartesian elements");
    if (h->prec == FLOU_HIP_FP64) {
        h->l64 = nd == 1 ? plain_f64_1d(n, h->eq, kvol, h->tp, h->curved)
                         : (nd == 2 ? plain_f64_2d(n, h->eq, kvol, h->tp, h->curved)
                                    : plain_f64_3d(n, h->eq, kvol, h->tp, h->curved));
        if (!h->l64) return bad(FLOU_HIP_EUNSUPPORTED, "no kernel for this combination");
    } else {
        h->l32 = nd == 1 ? plain_f32_1d(n, h->eq, kvol, h->tp, h->curved)
                         : (nd == 2 ? plain_f32_2d(n, h->eq, kvol, h->tp, h->curved)
                                    : plain_f32_3d(n, h->eq, kvol, h->tp, h->curved));
        if (!h->l32) return bad(FLOU_HIP_EUNSUPPORTED, "no kernel for this combination");
    }
    // adjoint kernels: the StdAverage surface average with LxF or no dissipation has a
    // hand-written face-flux VJP; every other Euler surface flux (Chandrasekhar average,
    // Scalar and Matrix dissipation) differentiates the plain flux in dual numbers (gen)
    const int gen = h->eq == FLOU_HIP_EULER &&
                    (h->savg != FLOU_HIP_AVG_STD || h->sdiss == FLOU_HIP_DISS_MATRIX ||
                     h->sdiss == FLOU_HIP_DISS_SCALAR);
    if (h->vol == FLOU_HIP_HYBRID) {
        h->vjp_why = "adjoint: the HybridDivOperator is not differentiated";
    } else if (kvol == VOL_SPLIT_NB) {
        h->vjp_why = "adjoint: the split form on Gauss nodes is not differentiated";
    } else if (h->prec == FLOU_HIP_FP64) {
        h->v64 = nd == 1 ? vjp_f64_1d(n, h->eq, h->vol, h->tp, h->curved, gen)
                         : (nd == 2 ? vjp_f64_2d(n, h->eq, h->vol, h->tp, h->curved, gen)
                                    : vjp_f64_3d(n, h->eq, h->vol, h->tp, h->curved, gen));
    } else {
        h->v32 = nd == 1 ? vjp_f32_1d(n, h->eq, h->vol, h->tp, h->curved, gen)
                         : (nd == 2 ? vjp_f32_2d(n, h->eq, h->vol, h->tp, h->curved, gen)
                                    : vjp_f32_3d(n, h->eq, h->vol, h->tp, h->curved, gen));
    }
    if (!h->v64 && !h->v32 && h->vjp_why.empty()) h->vjp_why = "adjoint: no kernel for this combination";

    // FP32 fast path: 2-D and 3-D split-form Euler with the StdAverage/LxF surface flux on
    // axis-aligned affine elements (diagonal metric, frames n = ±e_d) or on curved elements
    // (per-node metric and J, per-face-node frames), GLL traces
    // (l = e_1/e_n and ∂g supported on the end nodes up to 1e-12) and a vanishing D♯
    // diagonal, with fewer than 2^30 nodes (32-bit byte offsets per variable) and every face
    // orientation 0 (the fast kernels pair face nodes by index).  Anything else runs the
    // plain kernel in FP32.
    std::vector<float> fe, fds, fnj, fws, fiw;
    if (h->prec == FLOU_HIP_FP32 && (nd == 2 || nd == 3) && h->eq == FLOU_HIP_EULER &&
        h->vol == FLOU_HIP_SPLIT && h->savg == FLOU_HIP_AVG_STD &&
        (h->sdiss == FLOU_HIP_DISS_NONE || h->sdiss == FLOU_HIP_DISS_LXF) && n >= 4 &&
        h->ndofs < (int64_t(1) << 30) && !oriented) {
        bool ok = true;
        double dmax = 0.0;
        for (int i = 0; i < n * n; ++i) dmax = std::max(dmax, std::fabs(d->Dsharp[i]));
        for (int j = 0; j < n; ++j) {
            if (std::fabs(d->l_minus[j] - (j == 0 ? 1.0 : 0.0)) > 1e-12) ok = false;
            if (std::fabs(d->l_plus[j] - (j == n - 1 ? 1.0 : 0.0)) > 1e-12) ok = false;
            if (j > 0 && std::fabs(d->dg_minus[j]) > 1e-12 * std::fabs(d->dg_minus[0])) ok = false;
            if (j < n - 1 && std::fabs(d->dg_plus[j]) > 1e-12 * std::fabs(d->dg_plus[n - 1])) ok = false;
            if (std::fabs(d->Dsharp[j * n + j]) > 1e-10 * dmax) ok = false;
        }
        const int nd2 = nd * nd;
        // affine elements must be axis-aligned; curved ones carry their metric per node
        for (int64_t e = 0; e < ne && ok && !h->curved; ++e)
            for (int r = 0; r < nd; ++r)
                for (int c = 0; c < nd; ++c)
                    if (r != c && d->elem_metric[e * nd2 + r + nd * c] != 0.0) ok = false;
        for (int64_t f = 0; f < nf && ok && !h->curved; ++f) {
            const int dd = d->face_elempos[f * 2] >> 1;
            for (int c = 0; c < nd; ++c) {
                const double nc = d->face_frames[f * 3 * nd + c];
                if (c == dd ? std::fabs(nc) != 1.0 : nc != 0.0) ok = false;
            }
        }
        if (ok) h->lfast = fast_launcher(nd, n, h->tp, h->curved != 0);
        if (h->lfast && !h->curved) {
            // per element: Ja_cc/J (slots 0..nd-1), 1/J (slot 3), 1/Ja_cc (slots 4..3+nd)
            const int64_t cnt = uniform ? 1 : ne;
            fe.assign(cnt * 8, 0.0f);
            for (int64_t e = 0; e < cnt; ++e) {
                const double J = d->elem_jac[e];
                for (int c = 0; c < nd; ++c) {
                    const double ja = d->elem_metric[e * nd2 + c * (nd + 1)];
                    fe[e * 8 + c] = (float)(ja / J);        // Ja_cc / J
                    fe[e * 8 + 4 + c] = (float)(1.0 / ja);  // lift scale 1 / Ja_cc
                }
                fe[e * 8 + 3] = (float)(1.0 / J);
            }
            h->fegeom_stride = uniform ? 0 : 8;
        }
        if (h->lfast && h->curved) {
            // the curved kernel's face data: normal and surface J of each face node, the 4 of
            // the 3nd+1 values per node it uses, as one 16-byte vector
            const int64_t nq = nf * h->npf;
            fnj.assign(nq * 4, 0.0f);
            for (int64_t q = 0; q < nq; ++q) {
                for (int c = 0; c < nd; ++c) fnj[q * 4 + c] = (float)fg[q * FG + c];
                fnj[q * 4 + 3] = (float)fg[q * FG + 3 * nd];
            }
        }
        if (h->lfast && h->v32) {
            // the FP32 adjoint on this path (any of the three two-point fluxes):
            // S_ij = ω_i D♯_ij, antisymmetric off the zero diagonal (SBP on GLL nodes)
            double smax = 0.0, asym = 0.0;
            fws.resize(n * n);
            fiw.resize(n);
            for (int i = 0; i < n; ++i) {
                fiw[i] = (float)(1.0 / d->omega[i]);
                for (int j = 0; j < n; ++j) {
                    const double sij = d->omega[i] * d->Dsharp[i * n + j], sji = d->omega[j] * d->Dsharp[j * n + i];
                    smax = std::max(smax, std::fabs(sij));
                    asym = std::max(asym, std::fabs(sij + sji));
                    fws[i * n + j] = (float)(0.5 * (sij - sji));
                }
            }
            const bool walls = std::find(nb.begin(), nb.end(), -1) != nb.end();
            if (asym <= 1e-10 * smax) h->vfast = fast_vjp_launcher(nd, n, h->tp, h->curved != 0, walls);
        }
        if (h->lfast) {
            fds.resize(n * n);
            for (int i = 0; i < n * n; ++i) fds[i] = (float)d->Dsharp[i];
            h->lift_m = (float)d->dg_minus[0];
            h->lift_p = (float)d->dg_plus[n - 1];
            h->fdiss = h->sdiss == FLOU_HIP_DISS_LXF ? (float)h->intensity : 0.0f;
        }
    }

    // device uploads
    if (hipGetDevice(&h->device) != hipSuccess) return bad(FLOU_HIP_EDEVICE, "no HIP device");
    {
        // the allocation-free adjoint step's work memory on this device (a failure here only
        // disables that entry point)
        const hipError_t e = scratch_address(h->device, &h->d_scratch);
        if (e != hipSuccess) {
            h->d_scratch = nullptr;
            h->scratch_why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        }
    }
    if (h->lfast) {
        int rcf = fe.empty() ? 0 : upload(h, &h->d_fegeom, fe);
        rcf |= upload(h, &h->d_fdsharp, fds);
        if (!fnj.empty()) rcf |= upload(h, &h->d_fnj, fnj);
        if (h->vfast) rcf |= upload(h, &h->d_fwsh, fws) | upload(h, &h->d_fiw, fiw);
        if (rcf) return bad(FLOU_HIP_ENOMEM, "device upload failed: " + h->err);
    }
    int rc = 0;
    rc |= upload_real(h, &h->d_ops, ops);
    rc |= upload(h, (void**)&h->d_nb, nb);
    rc |= upload(h, (void**)&h->d_finfo, finfo);
    rc |= upload(h, (void**)&h->d_fgi, fgi);
    rc |= upload(h, (void**)&h->d_fid, fid);
    rc |= upload_real(h, &h->d_egeom, eg);
    rc |= upload_real(h, &h->d_fgeom, fg);
    rc |= upload_real(h, &h->d_ngeom, ng);
    rc |= upload(h, (void**)&h->d_dxe, dxe);
    if (h->eq == FLOU_HIP_EULER)
        rc |= upload(h, (void**)&h->d_evol, std::vector<double>(d->elem_volume, d->elem_volume + ne));
    rc |= upload(h, (void**)&h->d_jac, jac);
    rc |= upload(h, (void**)&h->d_w, w);
    h->bctype.assign(std::max(h->nbd, 1), FLOU_HIP_BC_OUTFLOW);
    h->bcconst.assign(std::max(h->nbd, 1) * h->nv, 0.0);
    rc |= upload(h, (void**)&h->d_bctype, h->bctype);
    rc |= upload_real(h, &h->d_bcconst, h->bcconst);
    if (rc) return bad(FLOU_HIP_ENOMEM, "device upload failed: " + h->err);
    if (hipMalloc((void**)&h->d_err, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&h->d_partial, reduce_max_blocks() * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&h->d_out, 4 * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void**)&h->h_out, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bad(FLOU_HIP_ENOMEM, "device allocation failed");
    if (hipMemset(h->d_err, 0, sizeof(int)) != hipSuccess) return bad(FLOU_HIP_EDEVICE, "hipMemset failed");
    *out = h;
    return FLOU_HIP_OK;
}

int flou_hip_destroy(flou_hip_handle* h) {
    if (!h) return FLOU_HIP_EINVAL;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    else (void)hipDeviceSynchronize();
    free_all(h);
    delete h;
    return FLOU_HIP_OK;
}

int flou_hip_set_stream(flou_hip_handle* h, void* stream) {
    if (!h) return FLOU_HIP_EINVAL;
    h->stream = (hipStream_t)stream;
    return FLOU_HIP_OK;
}

int flou_hip_set_option(flou_hip_handle* h, int32_t which, int32_t value) {
    if (!h) return FLOU_HIP_EINVAL;
    if (which == 0) { h->typo = value ? 1 : 0; return FLOU_HIP_OK; }
    if (which == 1) { h->force_plain = value ? 1 : 0; return FLOU_HIP_OK; }
    if (which == 2) return h->lfast ? 1 : 0;  // query: does FP32 run the fast kernel?
    if (which == 3) return h->vfast ? 1 : 0;  // query: does the FP32 adjoint run the fast kernel?
    return fail(h, FLOU_HIP_EINVAL, "unknown option");
}

int flou_hip_set_bc(flou_hip_handle* h, int32_t ib, int32_t type, const double* data) {
    if (!h) return FLOU_HIP_EINVAL;
    if (ib < 0 || ib >= h->nbd) return fail(h, FLOU_HIP_EINVAL, "boundary index out of range");
    if (type < FLOU_HIP_BC_DIRICHLET || type > FLOU_HIP_BC_SLIP) return fail(h, FLOU_HIP_EINVAL, "unknown BC type");
    if (h->eq == FLOU_HIP_ADVECTION && (type == FLOU_HIP_BC_SLIP))
        return fail(h, FLOU_HIP_EINVAL, "slip walls need the Euler equations");
    if ((type == FLOU_HIP_BC_DIRICHLET || type == FLOU_HIP_BC_INFLOW) && !data)
        return fail(h, FLOU_HIP_EINVAL, "BC data missing");
    h->bctype[ib] = type;
    if (type == FLOU_HIP_BC_INFLOW)
        for (int v = 0; v < h->nv; ++v) h->bcconst[ib * h->nv + v] = data[v];
    if (type == FLOU_HIP_BC_DIRICHLET) {
        // `data` covers every face; copy only the faces of this boundary
        const int64_t per = (int64_t)h->npf * h->nv;
        if (h->bcstate.empty()) h->bcstate.assign(h->nf * per, 0.0);
        std::vector<int> fid(h->ne * 2 * h->nd), finfo(h->ne * 2 * h->nd), nb(h->ne * 2 * h->nd);
        HIPCHECK(h, hipMemcpy(fid.data(), h->d_fid, fid.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(h, hipMemcpy(finfo.data(), h->d_finfo, finfo.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(h, hipMemcpy(nb.data(), h->d_nb, nb.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t s = 0; s < fid.size(); ++s)
            if (nb[s] < 0 && finfo_bd(finfo[s]) == ib)
                for (int64_t q = 0; q < per; ++q) h->bcstate[fid[s] * per + q] = data[fid[s] * per + q];
        if (h->d_bcstate) { (void)hipFree(h->d_bcstate); h->d_bcstate = nullptr; }
        if (upload_real(h, &h->d_bcstate, h->bcstate)) return FLOU_HIP_ENOMEM;
    }
    if (h->d_bctype) HIPCHECK(h, hipMemcpy(h->d_bctype, h->bctype.data(), h->bctype.size() * sizeof(int), hipMemcpyHostToDevice));
    if (h->d_bcconst) { (void)hipFree(h->d_bcconst); h->d_bcconst = nullptr; }
    if (upload_real(h, &h->d_bcconst, h->bcconst)) return FLOU_HIP_ENOMEM;
    return FLOU_HIP_OK;
}

int64_t flou_hip_ndofs(const flou_hip_handle* h) { return h ? h->ndofs : -1; }

size_t flou_hip_state_bytes(const flou_hip_handle* h) {
    return h ? (size_t)h->ndofs * h->nv * bytes_of(h) : 0;
}

int flou_hip_alloc(flou_hip_handle* h, void** dptr, size_t bytes) {
    if (!h || !dptr) return FLOU_HIP_EINVAL;
    HIPCHECK(h, hipMalloc(dptr, bytes));
    return FLOU_HIP_OK;
}

int flou_hip_free(flou_hip_handle* h, void* dptr) {
    if (!h) return FLOU_HIP_EINVAL;
    HIPCHECK(h, hipFree(dptr));
    return FLOU_HIP_OK;
}

int flou_hip_upload(flou_hip_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h) return FLOU_HIP_EINVAL;
    HIPCHECK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return FLOU_HIP_OK;
}

int flou_hip_download(flou_hip_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h) return FLOU_HIP_EINVAL;
    HIPCHECK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return FLOU_HIP_OK;
}

int flou_hip_synchronize(flou_hip_handle* h) {
    if (!h) return FLOU_HIP_EINVAL;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return FLOU_HIP_OK;
}

int flou_hip_rhs(flou_hip_handle* h, void* dQ, const void* Q, double t) {
    (void)t;  // the periodic/time-independent BCs of this path do not depend on t
    if (!h || !dQ || !Q) return FLOU_HIP_EINVAL;
    if (overlaps(dQ, Q, flou_hip_state_bytes(h))) return fail(h, FLOU_HIP_EINVAL, "dQ and Q overlap");
    return launch_stage(h, MODE_RHS, 0, 0.0, 0.0, 0.0, Q, nullptr, nullptr, dQ);
}

namespace {
// Zhang & Shu limiter in place on Q (Euler.jl:619-653), one thread per element.
int apply_limiter(flou_hip_handle* h, void* Q, double minval) {
    hipError_t e;
    if (h->prec == FLOU_HIP_FP64)
        e = launch_zhang_shu<double>((double*)Q, h->ndofs, h->ne, h->npe, h->nd, h->int_npe, h->d_jac, h->d_w,
                                     h->d_evol, minval, h->gamma, h->stream);
    else
        e = launch_zhang_shu<float>((float*)Q, h->ndofs, h->ne, h->npe, h->nd, h->int_npe, h->d_jac, h->d_w,
                                    h->d_evol, minval, h->gamma, h->stream);
    if (e != hipSuccess) return fail(h, FLOU_HIP_EDEVICE, std::string("limiter launch: ") + hipGetErrorString(e));
    return FLOU_HIP_OK;
}
}  // namespace

int flou_hip_set_limiter(flou_hip_handle* h, int32_t kind, double minval) {
    if (!h) return FLOU_HIP_EINVAL;
    if (kind == FLOU_HIP_LIMITER_NONE) {
        h->lim_kind = kind;
        return FLOU_HIP_OK;
    }
    if (kind != FLOU_HIP_LIMITER_ZHANG_SHU) return fail(h, FLOU_HIP_EINVAL, "unknown limiter");
    if (h->eq != FLOU_HIP_EULER) return fail(h, FLOU_HIP_EUNSUPPORTED, "the Zhang & Shu limiter needs the Euler equations");
    if (!std::isfinite(minval))
        return fail(h, FLOU_HIP_EINVAL, "The minimum value must be specified when using the limiter of Zhang & Shu.");
    h->lim_kind = kind;
    h->lim_min = minval;
    return FLOU_HIP_OK;
}

int flou_hip_limit(flou_hip_handle* h, void* Q, double minval) {
    if (!h || !Q) return FLOU_HIP_EINVAL;
    if (h->eq != FLOU_HIP_EULER) return fail(h, FLOU_HIP_EUNSUPPORTED, "the Zhang & Shu limiter needs the Euler equations");
    if (!std::isfinite(minval))
        return fail(h, FLOU_HIP_EINVAL, "The minimum value must be specified when using the limiter of Zhang & Shu.");
    return apply_limiter(h, Q, minval);
}

int flou_hip_monitor(flou_hip_handle* h, const void* Q, int32_t kind, double* out) {
    if (!h || !Q || !out) return FLOU_HIP_EINVAL;
    if (h->eq != FLOU_HIP_EULER) return fail(h, FLOU_HIP_EUNSUPPORTED, "the monitors need the Euler equations");
    if (kind != FLOU_HIP_MONITOR_KINETIC_ENERGY && kind != FLOU_HIP_MONITOR_ENTROPY)
        return fail(h, FLOU_HIP_EINVAL, "Unknown monitor");
    hipError_t e;
    if (h->prec == FLOU_HIP_FP64)
        e = launch_monitor<double>((const double*)Q, h->ndofs, h->int_npe, h->nd, h->gamma, kind, h->d_jac,
                                   h->d_w, h->d_partial, reduce_max_blocks(), h->d_out, h->d_err, h->stream);
    else
        e = launch_monitor<float>((const float*)Q, h->ndofs, h->int_npe, h->nd, h->gamma, kind, h->d_jac,
                                  h->d_w, h->d_partial, reduce_max_blocks(), h->d_out, h->d_err, h->stream);
    if (e != hipSuccess) return fail(h, FLOU_HIP_EDEVICE, hipGetErrorString(e));
    HIPCHECK(h, hipMemcpyAsync(h->h_out, h->d_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    *out = h->h_out[0];
    return flou_hip_check(h);
}

int flou_hip_equispaced(flou_hip_handle* h, int32_t m, const double* node2eq, const void* Q, int32_t ncols,
                        void* Qe) {
    if (!h || !node2eq || !Q || !Qe) return FLOU_HIP_EINVAL;
    if (m < 1) return fail(h, FLOU_HIP_EINVAL, "the number of equispaced points must be positive");
    if (ncols < 1) return fail(h, FLOU_HIP_EINVAL, "ncols must be positive");
    const size_t na = (size_t)m * h->n;
    for (size_t k = 0; k < na; ++k)
        if (!std::isfinite(node2eq[k])) return fail(h, FLOU_HIP_EINVAL, "node2eq must be finite");
    if (equispaced_lds_bytes(h->nd, h->n, m, bytes_of(h), nullptr) > 65536)
        return fail(h, FLOU_HIP_EUNSUPPORTED, "too many points per element for the equispaced projection");
    if (h->eqA.size() != na || !std::equal(h->eqA.begin(), h->eqA.end(), node2eq)) {
        // the previous upload may still be reading the host copy
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        if (h->eqA.size() < na || !h->d_eqA) {
            if (h->d_eqA) HIPCHECK(h, hipFree(h->d_eqA));
            h->d_eqA = nullptr;
            if (hipMalloc(&h->d_eqA, na * sizeof(double)) != hipSuccess) {
                h->eqA.clear();
                return fail(h, FLOU_HIP_ENOMEM, "device allocation failed");
            }
        }
        h->eqA.assign(node2eq, node2eq + na);
        HIPCHECK(h, hipMemcpyAsync(h->d_eqA, h->eqA.data(), na * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
    hipError_t e;
    if (h->prec == FLOU_HIP_FP64)
        e = launch_equispaced<double>((const double*)Q, (double*)Qe, h->d_eqA, h->nd, h->n, m, h->ne, ncols,
                                      h->stream);
    else
        e = launch_equispaced<float>((const float*)Q, (float*)Qe, h->d_eqA, h->nd, h->n, m, h->ne, ncols,
                                     h->stream);
    if (e != hipSuccess) return fail(h, FLOU_HIP_EDEVICE, hipGetErrorString(e));
    return FLOU_HIP_OK;
}

int flou_hip_rk_stage(flou_hip_handle* h, int32_t scheme, int32_t stage, double dt, double t,
                      const void* u_in, const void* u_base, void* reg, void* u_out) {
    (void)t;
    if (!h || !u_in || !u_out) return FLOU_HIP_EINVAL;
    const size_t bytes = flou_hip_state_bytes(h);
    if (overlaps(u_in, u_out, bytes)) return fail(h, FLOU_HIP_EINVAL, "u_out overlaps u_in");
    if (scheme == FLOU_HIP_SSPRK33) {
        if (stage < 0 || stage > 2) return fail(h, FLOU_HIP_EINVAL, "SSPRK33 has stages 0..2");
        if (stage > 0 && !u_base) return fail(h, FLOU_HIP_EINVAL, "u_base missing");
        // stage 2 may update u_base in place (u_out == u_base, point-wise); any other
        // overlap would let one workgroup read what another has written
        if (overlaps(u_base, u_out, bytes) && (stage == 1 || u_base != u_out))
            return fail(h, FLOU_HIP_EINVAL, "u_out overlaps u_base");
        if (overlaps(u_base, u_in, bytes) && u_base != u_in) return fail(h, FLOU_HIP_EINVAL, "u_base overlaps u_in");
        const int mode = stage == 0 ? MODE_SSP1 : (stage == 1 ? MODE_SSP2 : MODE_SSP3);
        int rc = launch_stage(h, mode, 0, dt, 0.0, 0.0, u_in, u_base, nullptr, u_out);
        if (!rc && h->lim_kind != FLOU_HIP_LIMITER_NONE) rc = apply_limiter(h, u_out, h->lim_min);
        return rc;
    }
    double A, B;
    if (!tableau(scheme, stage, A, B)) return fail(h, FLOU_HIP_EINVAL, "unknown scheme or stage");
    if (!reg) return fail(h, FLOU_HIP_EINVAL, "2N schemes need the register buffer");
    if (overlaps(reg, u_in, bytes) || overlaps(reg, u_out, bytes))
        return fail(h, FLOU_HIP_EINVAL, "register overlaps a state buffer");
    int rc = launch_stage(h, MODE_2N, stage == 0, dt, A, B, u_in, nullptr, reg, u_out);
    if (!rc && h->lim_kind != FLOU_HIP_LIMITER_NONE) rc = apply_limiter(h, u_out, h->lim_min);
    return rc;
}

int flou_hip_max_dt(flou_hip_handle* h, const void* Q, double cfl, double* out) {
    if (!h || !Q || !out) return FLOU_HIP_EINVAL;
    hipError_t e;
    if (h->prec == FLOU_HIP_FP64)
        e = launch_max_dt<double>((const double*)Q, h->ndofs, h->npe, h->nd, h->eq, h->gamma, h->anorm,
                                  cfl, h->d_dxe, h->d_partial, h->d_out, h->d_err, h->stream);
    else
        e = launch_max_dt<float>((const float*)Q, h->ndofs, h->npe, h->nd, h->eq, h->gamma, h->anorm,
                                 cfl, h->d_dxe, h->d_partial, h->d_out, h->d_err, h->stream);
    if (e != hipSuccess) return fail(h, FLOU_HIP_EDEVICE, hipGetErrorString(e));
    HIPCHECK(h, hipMemcpyAsync(h->h_out, h->d_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    *out = h->h_out[0];
    return FLOU_HIP_OK;
}

int flou_hip_integrate(flou_hip_handle* h, const void* Q, int32_t ivar, double* out) {
    if (!h || !Q || !out) return FLOU_HIP_EINVAL;
    if (ivar < 0 || ivar >= h->nv) return fail(h, FLOU_HIP_EINVAL, "ivar out of range");
    hipError_t e;
    if (h->prec == FLOU_HIP_FP64)
        e = launch_integrate<double>((const double*)Q + (int64_t)ivar * h->ndofs, h->ndofs, h->int_npe,
                                     h->d_jac, h->d_w, h->d_partial, h->d_out, h->stream);
    else
        e = launch_integrate<float>((const float*)Q + (int64_t)ivar * h->ndofs, h->ndofs, h->int_npe,
                                    h->d_jac, h->d_w, h->d_partial, h->d_out, h->stream);
    if (e != hipSuccess) return fail(h, FLOU_HIP_EDEVICE, hipGetErrorString(e));
    HIPCHECK(h, hipMemcpyAsync(h->h_out, h->d_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    *out = h->h_out[0];
    return FLOU_HIP_OK;
}

int flou_hip_check(flou_hip_handle* h) {
    if (!h) return FLOU_HIP_EINVAL;
    int flag = 0;
    HIPCHECK(h, hipMemcpyAsync(&flag, h->d_err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (flag) {
        HIPCHECK(h, hipMemsetAsync(h->d_err, 0, sizeof(int), h->stream));
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        return fail(h, FLOU_HIP_EDOMAIN,
                    (flag & ERR_SQRT_NEG)  ? "DomainError: sqrt of a negative number (non-physical state)"
                    : (flag & ERR_LOG_NEG) ? "DomainError: log of a negative number (non-physical state)"
                                           : "DomainError: power of a negative number (non-physical state)");
    }
    return FLOU_HIP_OK;
}

int flou_hip_timeintegrate(flou_hip_handle* h, int32_t scheme, void* Q, double dt, int64_t nsteps,
                           double cfl, double* t_inout) {
    if (!h || !Q || nsteps < 0) return FLOU_HIP_EINVAL;
    if (scheme < FLOU_HIP_SSPRK33 || scheme > FLOU_HIP_ORK256) return fail(h, FLOU_HIP_EINVAL, "unknown scheme");
    if (!(dt > 0.0)) return fail(h, FLOU_HIP_EINVAL, "dt must be positive");
    const size_t bytes = flou_hip_state_bytes(h);
    for (int i = 0; i < 2; ++i)
        if (!h->d_work[i]) HIPCHECK(h, hipMalloc(&h->d_work[i], bytes));
    double t = t_inout ? *t_inout : 0.0;
    // 2N schemes ping-pong the state between Q and work[0] (the register lives in
    // work[1]); `cur` carries over from step to step and is copied back once at the end.
    void* cur = Q;
    void* oth = h->d_work[0];
    int rc = 0;
    for (int64_t step = 0; step < nsteps && !rc; ++step) {
        double dts = dt;
        if (cfl > 0.0) {
            double m;
            rc = flou_hip_max_dt(h, cur, cfl, &m);
            if (rc) break;
            dts = std::min(m, dt);
        }
        if (scheme == FLOU_HIP_SSPRK33) {
            rc = flou_hip_rk_stage(h, scheme, 0, dts, t, Q, nullptr, nullptr, h->d_work[0]);
            if (!rc) rc = flou_hip_rk_stage(h, scheme, 1, dts, t, h->d_work[0], Q, nullptr, h->d_work[1]);
            if (!rc) rc = flou_hip_rk_stage(h, scheme, 2, dts, t, h->d_work[1], Q, nullptr, Q);
        } else {
            for (int s = 0; s < 5 && !rc; ++s) {
                rc = flou_hip_rk_stage(h, scheme, s, dts, t, cur, nullptr, h->d_work[1], oth);
                std::swap(cur, oth);
            }
        }
        t += dts;
    }
    if (rc) return rc;
    if (cur != Q) HIPCHECK(h, hipMemcpyAsync(Q, cur, bytes, hipMemcpyDeviceToDevice, h->stream));
    if (t_inout) *t_inout = t;
    return flou_hip_check(h);
}

int flou_hip_rhs_vjp(flou_hip_handle* h, const void* Q, const void* dQbar, void* Qbar) {
    if (!h || !Q || !dQbar || !Qbar) return FLOU_HIP_EINVAL;
    const size_t bytes = flou_hip_state_bytes(h);
    if (overlaps(Qbar, dQbar, bytes) || overlaps(Qbar, Q, bytes))
        return fail(h, FLOU_HIP_EINVAL, "Qbar must not overlap Q or dQbar");
    return launch_vjp(h, Q, dQbar, 1.0, Qbar, 1.0, nullptr, 0.0, Qbar, nullptr, 0.0, 0.0, 0.0, nullptr);
}

int flou_hip_rk_step_record(flou_hip_handle* h, int32_t scheme, double dt, double t,
                            void* const* stages, void* u_next) {
    if (!h || !stages || !u_next) return FLOU_HIP_EINVAL;
    if (h->lim_kind != FLOU_HIP_LIMITER_NONE)
        return fail(h, FLOU_HIP_EUNSUPPORTED, "adjoint: the stage limiter is not differentiated");
    const int S = scheme == FLOU_HIP_SSPRK33 ? 3 : 5;
    if (scheme < FLOU_HIP_SSPRK33 || scheme > FLOU_HIP_ORK256) return fail(h, FLOU_HIP_EINVAL, "unknown scheme");
    for (int s = 0; s < S; ++s)
        if (!stages[s]) return fail(h, FLOU_HIP_EINVAL, "stage buffer missing");
    int rc = 0;
    if (scheme == FLOU_HIP_SSPRK33) {
        rc = flou_hip_rk_stage(h, scheme, 0, dt, t, stages[0], nullptr, nullptr, stages[1]);
        if (!rc) rc = flou_hip_rk_stage(h, scheme, 1, dt, t, stages[1], stages[0], nullptr, stages[2]);
        if (!rc) rc = flou_hip_rk_stage(h, scheme, 2, dt, t, stages[2], stages[0], nullptr, u_next);
    } else {
        if (!h->d_work[0]) HIPCHECK(h, hipMalloc(&h->d_work[0], flou_hip_state_bytes(h)));
        for (int s = 0; s < S && !rc; ++s)
            rc = flou_hip_rk_stage(h, scheme, s, dt, t, stages[s], nullptr, h->d_work[0],
                                   s + 1 < S ? stages[s + 1] : u_next);
    }
    return rc;
}

namespace {
// Checks of the two adjoint-step entry points; S = number of stages.
int check_step_vjp(flou_hip_handle* h, int32_t scheme, const void* const* stages, void* ubar, int* S) {
    if (!h || !stages || !ubar) return FLOU_HIP_EINVAL;
    if (h->lim_kind != FLOU_HIP_LIMITER_NONE)
        return fail(h, FLOU_HIP_EUNSUPPORTED, "adjoint: the stage limiter is not differentiated");
    if (scheme < FLOU_HIP_SSPRK33 || scheme > FLOU_HIP_ORK256) return fail(h, FLOU_HIP_EINVAL, "unknown scheme");
    *S = scheme == FLOU_HIP_SSPRK33 ? 3 : 5;
    for (int s = 0; s < *S; ++s)
        if (!stages[s]) return fail(h, FLOU_HIP_EINVAL, "stage buffer missing");
    const size_t bytes = flou_hip_state_bytes(h);
    // every reverse stage reads stages[s] for neighbour traces while writing the result:
    // ubar must share no byte with a stage buffer
    for (int s = 0; s < *S; ++s)
        if (overlaps(ubar, stages[s], bytes)) return fail(h, FLOU_HIP_EINVAL, "ubar overlaps a stage buffer");
    return FLOU_HIP_OK;
}

// The reverse RK step over the work states X, Y2, Z (each one state, distinct from the
// stage buffers and ubar; nothing is read from them before it is written).
int step_vjp(flou_hip_handle* h, int32_t scheme, int S, double dt, const void* const* stages, void* ubar, void* X,
             void* Y2, void* Z) {
    int rc;
    if (scheme == FLOU_HIP_SSPRK33) {
        // u1 = u + dt L(u); u2 = (3u + u1 + dt L(u1))/4; u3 = (u + 2u2 + 2dt L(u2))/3
        // s=2: ū2 = ⅔ū3 + L'(u2)ᵀ(⅔dt ū3) -> X, acc = ⅓ū3 -> Z
        rc = launch_vjp(h, stages[2], ubar, 2.0 * dt / 3.0, ubar, 2.0 / 3.0, nullptr, 0.0, X,
                        ubar, 0.0, 0.0, 1.0 / 3.0, Z);
        // s=1: ū1 = ¼ū2 + L'(u1)ᵀ(¼dt ū2) -> Y2, acc += ¾ū2
        if (!rc) rc = launch_vjp(h, stages[1], X, dt / 4.0, X, 0.25, Z, 0.0, Y2, X, 0.0, 1.0, 0.75, Z);
        // s=0: ū = ū1 + acc + L'(u)ᵀ(dt ū1) -> ubar
        if (!rc) rc = launch_vjp(h, stages[0], Y2, dt, Y2, 1.0, Z, 1.0, ubar, nullptr, 0.0, 0.0, 0.0, nullptr);
        return rc;
    }
    // 2N: reg_s = A_s reg_{s-1} + dt L(u_s), u_{s+1} = u_s + B_s reg_s.  Reverse with
    // ḡ_s = B_s ū_{s+1} + A_{s+1} ḡ_{s+1}, ū_s = ū_{s+1} + L'(u_s)ᵀ(dt ḡ_s).
    double A[5], B[5];
    for (int s = 0; s < S; ++s) tableau(scheme, s, A[s], B[s]);
    void* G[2] = {Y2, Z};
    // s = S-1: ḡ = B ū (folded into the scale), result ū -> X, ḡ_{S-2} -> G[0]
    rc = launch_vjp(h, stages[S - 1], ubar, dt * B[S - 1], ubar, 1.0, nullptr, 0.0, X,
                    ubar, B[S - 2], 0.0, A[S - 1] * B[S - 1], G[0]);
    int cur = 0;
    for (int s = S - 2; s >= 1 && !rc; --s) {
        rc = launch_vjp(h, stages[s], G[cur], dt, X, 1.0, nullptr, 0.0, X, G[cur], B[s - 1], 0.0, A[s],
                        G[1 - cur]);
        cur = 1 - cur;
    }
    if (!rc) rc = launch_vjp(h, stages[0], G[cur], dt, X, 1.0, nullptr, 0.0, ubar, nullptr, 0.0, 0.0, 0.0, nullptr);
    return rc;
}
}  // namespace

int flou_hip_rk_step_vjp(flou_hip_handle* h, int32_t scheme, double dt, double t,
                         const void* const* stages, void* ubar) {
    (void)t;
    int S = 0;
    const int rc = check_step_vjp(h, scheme, stages, ubar, &S);
    if (rc) return rc;
    const size_t bytes = flou_hip_state_bytes(h);
    for (int i = 0; i < 3; ++i)
        if (!h->d_work[i]) HIPCHECK(h, hipMalloc(&h->d_work[i], bytes));
    return step_vjp(h, scheme, S, dt, stages, ubar, h->d_work[0], h->d_work[1], h->d_work[2]);
}

int flou_hip_rk_step_vjp_noalloc(flou_hip_handle* h, int32_t scheme, double dt, double t,
                                 const void* const* stages, void* ubar) {
    (void)t;
    int S = 0;
    const int rc = check_step_vjp(h, scheme, stages, ubar, &S);
    if (rc) return rc;
    if (!h->d_scratch) return fail(h, FLOU_HIP_EDEVICE, "no scratch memory on this device: " + h->scratch_why);
    const size_t bytes = flou_hip_state_bytes(h);
    const size_t slot = (bytes + 255) / 256 * 256;  // each work state 256-byte aligned
    if (3 * slot > (size_t)FLOU_HIP_NOALLOC_SCRATCH_BYTES)
        return fail(h, FLOU_HIP_EINVAL,
                    "rk_step_vjp_noalloc: three states of " + std::to_string(bytes) + " bytes exceed the " +
                        std::to_string((size_t)FLOU_HIP_NOALLOC_SCRATCH_BYTES) + "-byte scratch buffer");
    char* base = (char*)h->d_scratch;
    return step_vjp(h, scheme, S, dt, stages, ubar, base, base + slot, base + 2 * slot);
}

}  // extern "C"
