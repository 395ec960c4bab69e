// Projection of the nodal values of every element onto equispaced points, the device half
// of Flou's VTKHDF output (FlouBiz.pointdata2VTKHDF, src/FlouSpatial/IO.jl:78-121).
//
// The reference multiplies each element's values by node2eq = A ⊗ ... ⊗ A (the dense
// m^d x n^d matrix of StdQuad.jl:53 / StdHex.jl:61, applied in StdRegions.jl:232-238).
// Here the product is sum-factorised: one contraction per direction through LDS, d*m*n^d
// multiply-adds per element instead of m^d*n^d, so the kernel streams Q once and writes Qe
// once (HBM-bound).  Results agree with the dense product to rounding.
//
// Work unit = one column of one element (n^d contiguous values of Q, m^d of Qe: the columns
// of the state matrix are element-major, so Q and Qe are flat arrays of units).  A block
// stages several consecutive units with coalesced loads.  For m = n <= 8 the sizes are
// compile-time constants: each thread contracts whole lines with A held in registers, all
// contractions stay in LDS, and the block's units move between HBM and LDS as one
// contiguous run of 16-byte vectors each way; other sizes take a generic kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "reduce.hpp"

namespace flouhip {

namespace {

constexpr int kThreads = 256;

__host__ __device__ inline int ipow(int b, int e) {
    int r = 1;
    for (int k = 0; k < e; ++k) r *= b;
    return r;
}

// Generic sizes.  Block = `upb` consecutive units: stage them in LDS, contract direction
// 0, 1, .. with A (m x n, row-major), ping-ponging between two buffers of `cap` values per
// unit; the last contraction writes to HBM.  Stage tensors are x-fastest: dims < p have m
// points, dims >= p have n.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_equispaced(const T* __restrict__ Q, T* __restrict__ Qe,
                                                         const double* __restrict__ A, int nd, int n, int m,
                                                         int64_t units, int upb, int cap) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* sA = reinterpret_cast<T*>(smem);
    T* buf = sA + m * n;
    const int tid = threadIdx.x;
    for (int k = tid; k < m * n; k += blockDim.x) sA[k] = (T)A[k];
    const int64_t u0 = (int64_t)blockIdx.x * upb;
    const int nu = (int)((units - u0) < (int64_t)upb ? (units - u0) : (int64_t)upb);
    const int npe = ipow(n, nd), mpe = ipow(m, nd);
    const T* q = Q + u0 * npe;
    for (int k = tid; k < nu * npe; k += blockDim.x) {
        const int le = k / npe;
        buf[le * 2 * cap + (k - le * npe)] = q[k];
    }
    T* qe = Qe + u0 * mpe;
    for (int p = 0; p < nd; ++p) {
        __syncthreads();
        const int before = ipow(m, p), after = ipow(n, nd - 1 - p);
        const int outc = before * m * after;
        const int src = (p & 1) ? cap : 0, dst = (p & 1) ? 0 : cap;
        const bool last = p == nd - 1;
        for (int k = tid; k < nu * outc; k += blockDim.x) {
            const int le = k / outc, o = k - le * outc;
            const int ib = o % before, r = o / before;
            const int ip = r % m, ia = r / m;
            const T* in = buf + le * 2 * cap + src + ib + before * n * ia;
            const T* a = sA + ip * n;
            T acc = T(0);
            for (int j = 0; j < n; ++j) acc += a[j] * in[j * before];
            if (last)
                qe[(int64_t)le * mpe + o] = acc;
            else
                buf[le * 2 * cap + dst + o] = acc;
        }
    }
}

template <int B, int E>
struct Pow {
    static constexpr int v = B * Pow<B, E - 1>::v;
};
template <int B>
struct Pow<B, 0> {
    static constexpr int v = 1;
};

// units per block of the fixed-size kernel: about 8 KB staged per block (32 bytes per
// thread in flight), whatever the precision
template <typename T, int ND, int N>
struct FixedShape {
    static constexpr int NPE = Pow<N, ND>::v;
    static constexpr int UPB = NPE * (int)sizeof(T) >= 8192 ? 1 : 8192 / (NPE * (int)sizeof(T));
};

// One contraction of the fixed-size kernel: direction P of UPB units, each thread taking
// whole lines (N inputs -> N outputs) with the N x N matrix in registers.  The two LDS
// regions of UPB*NPE values each alternate as source and destination.
template <typename T, int ND, int N, int P, int UPB>
__device__ __forceinline__ void fixed_pass(const T (&a)[N][N], T* buf, int nu, int tid) {
    constexpr int NPE = Pow<N, ND>::v;
    constexpr int BEFORE = Pow<N, P>::v, AFTER = Pow<N, ND - 1 - P>::v;
    constexpr int LINES = BEFORE * AFTER;
    const T* src = buf + ((P & 1) ? UPB * NPE : 0);
    T* dst = buf + ((P & 1) ? 0 : UPB * NPE);
    for (int k = tid; k < nu * LINES; k += kThreads) {
        const int le = k / LINES, l = k - le * LINES;
        const int ib = l % BEFORE, ia = l / BEFORE;
        const int off = le * NPE + ib + BEFORE * N * ia;
        T v[N];
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = src[off + j * BEFORE];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) acc += a[i][j] * v[j];
            dst[off + i * BEFORE] = acc;
        }
    }
}

// Copy of `count` values between HBM and LDS (either direction), in 16-byte vectors when
// both ends are 16-byte aligned and count is a whole number of vectors, else value by value.
template <typename T>
__device__ __forceinline__ void block_copy(T* __restrict__ dst, const T* __restrict__ src, int count, int tid) {
    constexpr int VEC = 16 / (int)sizeof(T);
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0 && count % VEC == 0) {
        const float4* s4 = reinterpret_cast<const float4*>(src);
        float4* d4 = reinterpret_cast<float4*>(dst);
        for (int k = tid; k < count / VEC; k += kThreads) d4[k] = s4[k];
    } else {
        for (int k = tid; k < count; k += kThreads) dst[k] = src[k];
    }
}

template <typename T, int ND, int N>
__global__ __launch_bounds__(kThreads) void k_equispaced_fixed(const T* __restrict__ Q, T* __restrict__ Qe,
                                                               const double* __restrict__ A, int64_t units) {
    constexpr int NPE = FixedShape<T, ND, N>::NPE, UPB = FixedShape<T, ND, N>::UPB;
    __shared__ __align__(16) T buf[2 * UPB * NPE];
    const int tid = threadIdx.x;
    T a[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) a[i][j] = (T)A[i * N + j];
    const int64_t u0 = (int64_t)blockIdx.x * UPB;
    const int nu = (int)((units - u0) < (int64_t)UPB ? (units - u0) : (int64_t)UPB);
    block_copy(buf, Q + u0 * NPE, nu * NPE, tid);
    __syncthreads();
    fixed_pass<T, ND, N, 0, UPB>(a, buf, nu, tid);
    if constexpr (ND > 1) {
        __syncthreads();
        fixed_pass<T, ND, N, 1, UPB>(a, buf, nu, tid);
    }
    if constexpr (ND > 2) {
        __syncthreads();
        fixed_pass<T, ND, N, 2, UPB>(a, buf, nu, tid);
    }
    __syncthreads();
    block_copy(Qe + u0 * NPE, buf + (ND % 2 == 1 ? UPB * NPE : 0), nu * NPE, tid);
}

template <typename T, int ND, int N>
hipError_t launch_fixed(const T* Q, T* Qe, const double* A, int64_t units, hipStream_t s) {
    constexpr int UPB = FixedShape<T, ND, N>::UPB;
    const int64_t blocks = (units + UPB - 1) / UPB;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_equispaced_fixed<T, ND, N>), dim3((unsigned)blocks), dim3(kThreads), 0, s, Q, Qe, A,
                       units);
    return hipGetLastError();
}

template <typename T, int ND>
hipError_t dispatch_fixed(const T* Q, T* Qe, const double* A, int n, int64_t units, hipStream_t s, bool& done) {
    done = true;
    switch (n) {
        case 2: return launch_fixed<T, ND, 2>(Q, Qe, A, units, s);
        case 3: return launch_fixed<T, ND, 3>(Q, Qe, A, units, s);
        case 4: return launch_fixed<T, ND, 4>(Q, Qe, A, units, s);
        case 5: return launch_fixed<T, ND, 5>(Q, Qe, A, units, s);
        case 6: return launch_fixed<T, ND, 6>(Q, Qe, A, units, s);
        case 7: return launch_fixed<T, ND, 7>(Q, Qe, A, units, s);
        case 8: return launch_fixed<T, ND, 8>(Q, Qe, A, units, s);
        default: done = false; return hipSuccess;
    }
}

}  // namespace

int equispaced_lds_bytes(int nd, int n, int m, int elem_bytes, int* upb_out) {
    const int big = n > m ? n : m;
    const int cap = ipow(big, nd);
    int upb = 1024 / cap;
    if (upb < 1) upb = 1;
    const int fixed = m * n * elem_bytes;
    while (upb > 1 && fixed + upb * 2 * cap * elem_bytes > 32768) --upb;
    if (upb_out) *upb_out = upb;
    return fixed + upb * 2 * cap * elem_bytes;
}

template <typename T>
hipError_t launch_equispaced(const T* Q, T* Qe, const double* A, int nd, int n, int m, int64_t ne, int ncols,
                             hipStream_t s) {
    if (ne <= 0 || ncols <= 0) return hipSuccess;
    if (nd < 1 || nd > 3 || n < 1 || m < 1) return hipErrorInvalidValue;
    const int64_t units = ne * ncols;
    if (m == n) {
        bool done = false;
        hipError_t e = nd == 1   ? dispatch_fixed<T, 1>(Q, Qe, A, n, units, s, done)
                       : nd == 2 ? dispatch_fixed<T, 2>(Q, Qe, A, n, units, s, done)
                                 : dispatch_fixed<T, 3>(Q, Qe, A, n, units, s, done);
        if (done) return e;
    }
    int upb = 1;
    const int lds = equispaced_lds_bytes(nd, n, m, (int)sizeof(T), &upb);
    if (lds > 65536) return hipErrorInvalidValue;
    const int cap = ipow(n > m ? n : m, nd);
    const int64_t blocks = (units + upb - 1) / upb;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_equispaced<T>, dim3((unsigned)blocks), dim3(kThreads), lds, s, Q, Qe, A, nd, n, m, units,
                       upb, cap);
    return hipGetLastError();
}

template hipError_t launch_equispaced<double>(const double*, double*, const double*, int, int, int, int64_t, int,
                                              hipStream_t);
template hipError_t launch_equispaced<float>(const float*, float*, const double*, int, int, int, int64_t, int,
                                             hipStream_t);

}  // namespace flouhip
