// Argument block and launchers of the FP32 fast element kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flouhip {

struct FastArgs {
    const float* Q;     // stage input, (ndofs, 5) column-major, ndofs < 2^30
    float* out;
    const float* ubase;
    float* reg;
    float dt, A, B;
    int mode, first;
    int64_t ndofs;
    int ne;
    const int* nb;      // ne*6 neighbour element (-1: boundary)
    const int* finfo;   // ne*6 (see common.hpp)
    const int* fgi;     // ne*6 index into fgeom
    const int* fid;     // ne*6 mesh face index
    const float* egeom; // per element: Ja_xx/J, Ja_yy/J, Ja_zz/J, 1/J, 1/Ja_xx, 1/Ja_yy, 1/Ja_zz, -
    int egeom_stride;   // 0 when every element is alike, else 8
    const float* fgeom; // per entry: frame (n, t, b) then surface J (10 floats)
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

// 3-D Euler split form, n = 4..8 nodes per direction, tp = two-point flux.
FastLauncher fast3d_launcher(int n, int tp);

}  // namespace flouhip
