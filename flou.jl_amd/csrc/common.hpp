// Shared host/device definitions of libflouhip: kernel argument blocks, tensor-product
// index helpers and launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flouhip {

enum { EQ_ADVECTION = 0, EQ_EULER = 1 };
enum { VOL_STRONG = 0, VOL_SPLIT = 1, VOL_HYBRID = 2 };
// Kernel-internal variants on nodes without boundaries (Gauss): the split form, whose
// surface terms use entropy-projected end states (OpDivergence.jl:300-436), and the hybrid
// form, whose subcell fluxes telescope from those terms (OpDivergence.jl:633-779).
enum { VOL_SPLIT_NB = 3, VOL_HYBRID_NB = 4 };
enum { TP_STDAVG = 0, TP_CHANDRASEKHAR = 1, TP_KENNEDY_GRUBER = 2 };
enum { AVG_STD = 0, AVG_CHANDRASEKHAR = 1 };
enum { DISS_NONE = 0, DISS_LXF = 1, DISS_MATRIX = 2, DISS_SCALAR = 3 };
enum { BC_DIRICHLET = 0, BC_INFLOW = 1, BC_OUTFLOW = 2, BC_SLIP = 3 };

// What the fused element kernel does with L(u_in) once it has it.
enum {
    MODE_RHS = 0,   // out = L(u_in)
    MODE_2N = 1,    // reg = A*reg + dt*L (reg = dt*L when first); out = u_in + B*reg
    MODE_SSP1 = 2,  // out = u_in + dt*L
    MODE_SSP2 = 3,  // out = ((3*ubase + u_in) + dt*L) / 4
    MODE_SSP3 = 4   // out = ((ubase + 2*u_in) + (2*dt)*L) / 3
};

// Error flags raised by kernels (the reference throws DomainError at these points).
enum { ERR_SQRT_NEG = 1, ERR_LOG_NEG = 2, ERR_POW_NEG = 4 };

// Packed 1-D operators, offsets in units of scalars.
__host__ __device__ constexpr int ops_D(int) { return 0; }
__host__ __device__ constexpr int ops_Ds(int n) { return n * n; }
__host__ __device__ constexpr int ops_Dsh(int n) { return 2 * n * n; }
__host__ __device__ constexpr int ops_lm(int n) { return 3 * n * n; }
__host__ __device__ constexpr int ops_lp(int n) { return 3 * n * n + n; }
__host__ __device__ constexpr int ops_gm(int n) { return 3 * n * n + 2 * n; }
__host__ __device__ constexpr int ops_gp(int n) { return 3 * n * n + 3 * n; }
__host__ __device__ constexpr int ops_w(int n) { return 3 * n * n + 4 * n; }
__host__ __device__ constexpr int ops_size(int n) { return 3 * n * n + 5 * n; }

// Per element-face table entry "finfo":
//   bit 0       : side of this element on the face (0 = master / side 1 in Flou, 1 = slave)
//   bits 1..3   : face position of the neighbour (0-based) when the face is interior
//   bits 4..27  : boundary index + 1 for boundary faces, 0 for interior faces
//   bits 28..30 : orientation of an interior face (mesh.faces[i].orientation)
__host__ __device__ inline int finfo_side(int x) { return x & 1; }
__host__ __device__ inline int finfo_nbpos(int x) { return (x >> 1) & 7; }
__host__ __device__ inline int finfo_bd(int x) { return ((x >> 4) & 0xffffff) - 1; }
__host__ __device__ inline int finfo_orient(int x) { return (x >> 28) & 7; }

template <typename T>
struct RhsArgs {
    const T* Q;       // stage input u_in, (ndofs, nv) column-major
    T* out;           // L(u_in) or the stage output
    const T* ubase;   // SSPRK33 base state
    T* reg;           // 2N register
    T dt, A, B;
    int mode, first;
    int64_t ndofs;
    int ne;
    const int* nb;     // ne*2nd neighbour element (-1: boundary)
    const int* finfo;  // ne*2nd, see above
    const int* fgi;    // ne*2nd index into fgeom
    const int* fid;    // ne*2nd mesh face index
    const T* egeom;    // per element: metric (nd*nd, column-major) then J
    int egeom_stride;  // 0 when every element has the same geometry
    const T* ngeom;    // curved: per node, component-major [nd*nd + 1][ndofs] (metric, J)
    const T* fgeom;    // per entry: frame (n, t, b: 3*nd) then surface J; curved: per face
                       // node, entry fgi*npf + f
    const T* ops;      // packed 1-D operators
    const T* bcstate;  // Dirichlet exterior states, nfaces * npf * nv
    const T* bcconst;  // inflow states, nboundaries * nv
    const int* bctype; // nboundaries
    int gather;        // 1 when l = (e_1, e_n) exactly: traces are node gathers
    int lift_end;      // 1 when ∂g± vanish exactly off the end nodes (GLL DGSEM lift)
    T gamma, intensity;
    T a0, a1, a2;      // advection velocity
    int savg, sdiss, typo;
    int* errflag;
    T blend;           // HybridDivOperator: blending constant c
    T sjac[3];         // HybridDivOperator: subcell-interface surface Jacobian per direction
};

// Block index -> logical block, so that consecutive logical blocks (neighbouring
// elements that share faces) land on the same XCD and its L2.  Workgroups are dispatched
// round-robin over the 8 XCDs of MI355X.
__device__ inline int xcd_remap(int bid, int nblocks) {
    const int q = nblocks >> 3, r = nblocks & 7;
    const int xcd = bid & 7, local = bid >> 3;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

template <int ND, int N>
struct Cfg {
    static constexpr int NPE = ND == 1 ? N : (ND == 2 ? N * N : N * N * N);
    static constexpr int NPF = ND == 1 ? 1 : (ND == 2 ? N : N * N);
    static constexpr int NF = 2 * ND;
    static constexpr int EPB = NPE >= 256 ? 1 : 256 / NPE;  // elements per block
    static constexpr int TPB = EPB * NPE;                   // threads per block
};

// Node index of position `p` along line `f` of direction `dir` (tpdofs, StdQuad.jl:116-124,
// StdHex.jl:135-145: lines enumerated with the lower transverse index fastest).
template <int ND, int N>
__host__ __device__ inline int line_node(int dir, int f, int p) {
    if (ND == 1) return p;
    if (ND == 2) return dir == 0 ? p + N * f : f + N * p;
    const int fi = f % N, fj = f / N;
    if (dir == 0) return p + N * (fi + N * fj);
    if (dir == 1) return fi + N * (p + N * fj);
    return fi + N * (fj + N * p);
}

// Face-node maps between the two sides of a face with orientation o, for face nodes
// numbered as line_node's lines (f = a + N*b on a quad face, a fastest): master2slave and
// slave2master of StdSegment.jl:153-167 (o = 1 reverses the segment) and StdQuad.jl:
// 138-181 (o = 1..3 rotate, 4..7 reflect the square).
template <int ND, int N>
__host__ __device__ inline int face_m2s(int f, int o) {
    if (ND < 2 || o == 0) return f;
    if (ND == 2) return N - 1 - f;
    const int a = f % N, b = f / N, r = N - 1;
    switch (o) {
        case 1: return b + N * (r - a);
        case 2: return (r - a) + N * (r - b);
        case 3: return (r - b) + N * a;
        case 4: return b + N * a;
        case 5: return (r - a) + N * b;
        case 6: return (r - b) + N * (r - a);
        default: return a + N * (r - b);
    }
}
template <int ND, int N>
__host__ __device__ inline int face_s2m(int f, int o) {
    if (ND < 2 || o == 0) return f;
    if (ND == 2) return N - 1 - f;
    const int a = f % N, b = f / N, r = N - 1;
    switch (o) {
        case 1: return (r - b) + N * a;
        case 2: return (r - a) + N * (r - b);
        case 3: return b + N * (r - a);
        case 4: return b + N * a;
        case 5: return (r - a) + N * b;
        case 6: return (r - b) + N * (r - a);
        default: return a + N * (r - b);
    }
}
// For own face node f of an interior face (finfo `info`): the neighbour's face node that
// meets it, and the master's index of the pair, which indexes the per-face-node geometry.
template <int ND, int N>
__host__ __device__ inline int face_nb_node(int f, int info) {
    const int o = finfo_orient(info);
    if (o == 0) return f;
    return finfo_side(info) ? face_s2m<ND, N>(f, o) : face_m2s<ND, N>(f, o);
}
template <int ND, int N>
__host__ __device__ inline int face_master_node(int f, int info) {
    const int o = finfo_orient(info);
    return (o == 0 || !finfo_side(info)) ? f : face_s2m<ND, N>(f, o);
}

// (line index, position) of `node` in direction `dir`.
template <int ND, int N>
__host__ __device__ inline void node_line(int node, int dir, int& f, int& p) {
    const int ix = node % N;
    const int iy = ND >= 2 ? (node / N) % N : 0;
    const int iz = ND >= 3 ? node / (N * N) : 0;
    if (ND == 1) { f = 0; p = ix; return; }
    if (ND == 2) {
        if (dir == 0) { f = iy; p = ix; } else { f = ix; p = iy; }
        return;
    }
    if (dir == 0) { f = iy + N * iz; p = ix; }
    else if (dir == 1) { f = ix + N * iz; p = iy; }
    else { f = ix + N * iy; p = iz; }
}

// Adjoint pass (vjp_plain.hip): V = (d out/dQ)^T (c·Y) at r.Q, then
//   o0 = a0·P0 + a1·P1 + V,  o1 = b0·o0 + b1·P1 + b2·P2   (null P*: term absent; o1 optional)
template <typename T>
struct VjpArgs {
    RhsArgs<T> r;
    const T* Y;
    T c;
    const T* P0;
    const T* P1;
    const T* P2;
    T a0, a1, b0, b1, b2;
    T* o0;
    T* o1;
};

typedef hipError_t (*RhsLauncherF64)(const RhsArgs<double>&, hipStream_t);
typedef hipError_t (*RhsLauncherF32)(const RhsArgs<float>&, hipStream_t);

// Generic ("plain") element kernels, one translation unit per (precision, dimension).
RhsLauncherF64 plain_f64_1d(int n, int eq, int vol, int tp, int curved);
RhsLauncherF64 plain_f64_2d(int n, int eq, int vol, int tp, int curved);
RhsLauncherF64 plain_f64_3d(int n, int eq, int vol, int tp, int curved);
RhsLauncherF32 plain_f32_1d(int n, int eq, int vol, int tp, int curved);
RhsLauncherF32 plain_f32_2d(int n, int eq, int vol, int tp, int curved);
RhsLauncherF32 plain_f32_3d(int n, int eq, int vol, int tp, int curved);

typedef hipError_t (*VjpLauncherF64)(const VjpArgs<double>&, hipStream_t);
typedef hipError_t (*VjpLauncherF32)(const VjpArgs<float>&, hipStream_t);
VjpLauncherF64 vjp_f64_1d(int n, int eq, int vol, int tp, int curved, int gen);
VjpLauncherF64 vjp_f64_2d(int n, int eq, int vol, int tp, int curved, int gen);
VjpLauncherF64 vjp_f64_3d(int n, int eq, int vol, int tp, int curved, int gen);
VjpLauncherF32 vjp_f32_1d(int n, int eq, int vol, int tp, int curved, int gen);
VjpLauncherF32 vjp_f32_2d(int n, int eq, int vol, int tp, int curved, int gen);
VjpLauncherF32 vjp_f32_3d(int n, int eq, int vol, int tp, int curved, int gen);

}  // namespace flouhip
