// FP64 reductions over the state: CFL time step (get_max_dt) and conservation integral
// (integrate), plus the small element-wise helpers the host API needs.
//
// Both reductions read the state once (HBM-bound) and reduce in two passes: a wavefront
// reduction with DPP/shuffles inside each block, one partial per block, then a single block
// over the partials.  The min is exact in any order; the sum is a fixed tree, so results
// are deterministic run to run.
#include "common.hpp"
#include "reduce.hpp"

namespace flouhip {

static constexpr int RB = 256;

__device__ inline double warp_min(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(x, o, 64);
        x = (y < x) ? y : x;
    }
    return x;
}
__device__ inline double warp_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

template <bool MIN>
__device__ inline double block_reduce(double x) {
    __shared__ double s[RB / 64];
    x = MIN ? warp_min(x) : warp_sum(x);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) s[w] = x;
    __syncthreads();
    if (w == 0) {
        x = (l < RB / 64) ? s[l] : (MIN ? INFINITY : 0.0);
        x = MIN ? warp_min(x) : warp_sum(x);
    }
    return x;
}

// Local Δt = cfl Δx / (|u| + c) per node in FP64 (FlouCommon/Euler.jl:116-135;
// LinearAdvection.jl:46-48), Δx = (vol / n^d)^(1/d) per element, min over nodes.
template <typename T>
__global__ void __launch_bounds__(RB) k_max_dt(const T* __restrict__ Q, int64_t ndofs, int npe,
                                               int nd, int eq, double gamma, double anorm,
                                               double cfl, const double* __restrict__ dxe,
                                               double* __restrict__ partial, int* err) {
    double m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < ndofs; i += (int64_t)gridDim.x * RB) {
        const double dx = dxe[i / npe];
        double loc;
        if (eq == EQ_ADVECTION) {
            loc = cfl * dx / anorm;
        } else {
            double q[5];
            for (int v = 0; v < nd + 2; ++v) q[v] = (double)Q[v * ndofs + i];
            double s = q[1] * q[1];
            for (int k = 1; k < nd; ++k) s = s + q[1 + k] * q[1 + k];
            const double p = (gamma - 1.0) * (q[nd + 1] - s / (2.0 * q[0]));
            const double c2 = gamma * p / q[0];
            if (c2 < 0.0) atomicOr(err, ERR_SQRT_NEG);
            const double c = sqrt(c2);
            double um;
            if (nd == 1) {
                um = fabs(q[1] / q[0]);
            } else {
                double u = q[1] / q[0];
                double ss = u * u;
                for (int k = 1; k < nd; ++k) {
                    u = q[1 + k] / q[0];
                    ss = ss + u * u;
                }
                um = sqrt(ss);
            }
            loc = cfl * dx / (um + c);
        }
        m = (loc < m) ? loc : m;
    }
    m = block_reduce<true>(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

// Σ_e Σ_i J_e ω_i Q_{i,ivar} (MultielementDiscontinuous.jl:148-160)
template <typename T>
__global__ void __launch_bounds__(RB) k_integrate(const T* __restrict__ Q, int64_t ndofs, int npe,
                                                  const double* __restrict__ jac,
                                                  const double* __restrict__ w,
                                                  double* __restrict__ partial) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < ndofs; i += (int64_t)gridDim.x * RB) {
        const int64_t e = i / npe;
        s += (jac[e] * w[i - e * npe]) * (double)Q[i];
    }
    s = block_reduce<false>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <bool MIN>
__global__ void __launch_bounds__(RB) k_finish(const double* __restrict__ partial, int n,
                                               double* __restrict__ out) {
    double x = MIN ? INFINITY : 0.0;
    for (int i = threadIdx.x; i < n; i += RB) {
        const double y = partial[i];
        x = MIN ? ((y < x) ? y : x) : x + y;
    }
    x = block_reduce<MIN>(x);
    if (threadIdx.x == 0) *out = x;
}

// Copy with conversion (FP64 snapshots of an FP32 state and back).
template <typename S, typename D>
__global__ void __launch_bounds__(RB) k_convert(const S* __restrict__ src, D* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
        dst[i] = (D)src[i];
}

static int grid_for(int64_t n) {
    int64_t g = (n + RB - 1) / RB;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (int)g;
}

template <typename T>
hipError_t launch_max_dt(const T* Q, int64_t ndofs, int npe, int nd, int eq, double gamma,
                         double anorm, double cfl, const double* dxe, double* partial,
                         double* out, int* err, hipStream_t s) {
    const int g = grid_for(ndofs);
    hipLaunchKernelGGL(k_max_dt<T>, dim3(g), dim3(RB), 0, s, Q, ndofs, npe, nd, eq, gamma, anorm,
                       cfl, dxe, partial, err);
    hipLaunchKernelGGL(k_finish<true>, dim3(1), dim3(RB), 0, s, partial, g, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_integrate(const T* Q, int64_t ndofs, int npe, const double* jac,
                            const double* w, double* partial, double* out, hipStream_t s) {
    const int g = grid_for(ndofs);
    hipLaunchKernelGGL(k_integrate<T>, dim3(g), dim3(RB), 0, s, Q, ndofs, npe, jac, w, partial);
    hipLaunchKernelGGL(k_finish<false>, dim3(1), dim3(RB), 0, s, partial, g, out);
    return hipGetLastError();
}

template <typename S, typename D>
hipError_t launch_convert(const S* src, D* dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<S, D>), dim3(grid_for(n)), dim3(RB), 0, s, src, dst, n);
    return hipGetLastError();
}

template hipError_t launch_max_dt<double>(const double*, int64_t, int, int, int, double, double,
                                          double, const double*, double*, double*, int*, hipStream_t);
template hipError_t launch_max_dt<float>(const float*, int64_t, int, int, int, double, double,
                                         double, const double*, double*, double*, int*, hipStream_t);
template hipError_t launch_integrate<double>(const double*, int64_t, int, const double*,
                                             const double*, double*, double*, hipStream_t);
template hipError_t launch_integrate<float>(const float*, int64_t, int, const double*,
                                            const double*, double*, double*, hipStream_t);
template hipError_t launch_convert<float, double>(const float*, double*, int64_t, hipStream_t);
template hipError_t launch_convert<double, float>(const double*, float*, int64_t, hipStream_t);

int reduce_max_blocks() { return 4096; }

}  // namespace flouhip
