// Argument block and launchers of the FP32 fast element kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flouhip {

struct FastArgs {
    const float* Q;     // stage input, (ndofs, nd + 2) column-major, ndofs < 2^30
    float* out;
    const float* ubase;
    float* reg;
    float dt, A, B;
    int mode, first;
    int64_t ndofs;
    int ne;
    const int* nb;      // ne*2nd neighbour element (-1: boundary)
    const int* finfo;   // ne*2nd (see common.hpp)
    const int* fgi;     // ne*2nd index into fgeom
    const int* fid;     // ne*2nd mesh face index
    const float* egeom; // per element: Ja_dd/J (d < nd, slots 0..2), 1/J (slot 3), 1/Ja_dd (slots 4..6)
    int egeom_stride;   // 0 when every element is alike, else 8
    const float* fgeom; // per entry: frame (n, t, b) then surface J (3nd + 1 floats); curved: per face node
    const float* ngeom; // curved: per node, component-major [nd*nd + 1][ndofs] (metric, J)
    const float4* fnj;  // curved: per face node of each face entry, (normal, surface J)
    const float* dsharp;// D♯, row-major n x n
    float lift_m, lift_p; // ∂g_minus[0], ∂g_plus[n-1]
    const float* bcstate;
    const float* bcconst;
    const int* bctype;
    float gamma, inv_gm1;
    float diss;         // LxF intensity (0 for the StdAverage surface flux)
    int* errflag;
};

typedef hipError_t (*FastLauncher)(const FastArgs&, hipStream_t);

// 2-D and 3-D Euler split form, n = 4..8 nodes per direction, tp = two-point flux, on
// axis-aligned affine elements or (curved) with per-node metric and J.
FastLauncher fast_launcher(int nd, int n, int tp, bool curved);

// FP32 adjoint pass on the fast path, the same contract as VjpArgs (common.hpp):
//   o0 = a0·P0 + a1·P1 + (∂out/∂Q)^T(c·Y) at f.Q,  o1 = b0·o0 + b1·P1 + b2·P2 (o1 optional),
// with out the rhs (f.Q, geometry, face tables, lift, γ and the LxF intensity as in FastArgs;
// f.out, f.ubase, f.reg, f.dt, f.A, f.B, f.mode and f.first unused).
struct FastVjpArgs {
    FastArgs f;
    const float* Y;
    float c;
    const float* P0;
    const float* P1;
    const float* P2;
    float a0, a1, b0, b1, b2;
    float* o0;
    float* o1;
    const float* wsh;  // S_ij = ω_i D♯_ij, row-major n x n (antisymmetric on GLL nodes)
    const float* iw;   // 1/ω_p
};

typedef hipError_t (*FastVjpLauncher)(const FastVjpArgs&, hipStream_t);

namespace fast {
// the kernels of one two-point flux (one translation unit each, which instantiates its own)
template <int TP>
FastVjpLauncher fvjp_tp(int nd, int n, bool curved, bool walls);
extern template FastVjpLauncher fvjp_tp<0>(int, int, bool, bool);
extern template FastVjpLauncher fvjp_tp<1>(int, int, bool, bool);
extern template FastVjpLauncher fvjp_tp<2>(int, int, bool, bool);
}  // namespace fast

// 2-D and 3-D Euler split form with the StdAverage (tp 0), Chandrasekhar (1) or
// Kennedy-Gruber (2) two-point flux, n = 4..8, on the fast path's elements (curved, or
// axis-aligned affine); walls: the mesh has boundary faces; nullptr otherwise.
inline FastVjpLauncher fast_vjp_launcher(int nd, int n, int tp, bool curved, bool walls) {
    switch (tp) {
        case 0: return fast::fvjp_tp<0>(nd, n, curved, walls);
        case 1: return fast::fvjp_tp<1>(nd, n, curved, walls);
        case 2: return fast::fvjp_tp<2>(nd, n, curved, walls);
        default: return nullptr;
    }
}

}  // namespace flouhip
