// Stage limiter and monitors of the Euler path (Equations/Euler.jl:541-654).
//
// k_zhang_shu: Flou's Zhang & Shu positivity limiter, one thread per element.  Element
// means, minima and the two pulls are sequential folds over the element's nodes (the
// reference's order), so the FP64 result is reproducible bit for bit; an element's nodes
// are contiguous per variable, so each thread streams its own cache lines.
// k_monitor: Σ Jω · s(Q) with s the kinetic energy or the mathematical entropy, FP64
// accumulation, block partials + one finishing block (fixed tree, deterministic).
#include "common.hpp"
#include "reduce.hpp"

namespace flouhip {

static constexpr int LB = 64;   // limiter: elements per block
static constexpr int MB = 256;  // monitor block

template <typename T, int NV>
__global__ void __launch_bounds__(LB) k_zhang_shu(T* __restrict__ Q, int64_t ndofs, int64_t ne, int npe,
                                                 int wnpe, const double* __restrict__ jac,
                                                 const double* __restrict__ w,
                                                 const double* __restrict__ vol, T minval, T gamma) {
    constexpr int ND = NV - 2;
    const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
    if (e >= ne) return;
    T* q[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = Q + v * ndofs + e * npe;
    auto jw = [&](int i) -> T {
        const int64_t g = e * npe + i;
        const int64_t k = g / wnpe;
        return (T)(jac[k] * w[g - k * wnpe]);
    };
    const T V = (T)vol[e];
    // element means and density minimum
    T qm[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) qm[v] = T(0);
    T rmin = q[0][0];
    for (int i = 0; i < npe; ++i) {
        const T c = jw(i);
#pragma unroll
        for (int v = 0; v < NV; ++v) qm[v] = qm[v] + c * q[v][i];
        rmin = q[0][i] < rmin ? q[0][i] : rmin;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) qm[v] = qm[v] / V;
    // density limiting
    T m = minval < qm[0] ? minval : qm[0];
    T th = fabs((qm[0] - m) / (qm[0] - rmin));
    if (th <= T(1))
        for (int i = 0; i < npe; ++i) q[0][i] = th * (q[0][i] - qm[0]) + qm[0];
    // pressure of the updated nodes: mean and minimum
    T pm = T(0), pmin = T(0);
    for (int i = 0; i < npe; ++i) {
        T s = q[1][i] * q[1][i];
#pragma unroll
        for (int k = 1; k < ND; ++k) s = s + q[1 + k][i] * q[1 + k][i];
        const T p = (gamma - T(1)) * (q[ND + 1][i] - s / (T(2) * q[0][i]));
        pm = pm + jw(i) * p;
        pmin = (i == 0 || p < pmin) ? p : pmin;
    }
    pm = pm / V;
    m = minval < pm ? minval : pm;
    th = fabs((pm - m) / (pm - pmin));
    if (th <= T(1))
        for (int i = 0; i < npe; ++i)
#pragma unroll
            for (int v = 0; v < NV; ++v) q[v][i] = th * (q[v][i] - qm[v]) + qm[v];
}

__device__ inline double mwarp_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

template <typename T>
__global__ void __launch_bounds__(MB) k_monitor(const T* __restrict__ Q, int64_t ndofs, int wnpe, int nd,
                                               double gamma, int kind, const double* __restrict__ jac,
                                               const double* __restrict__ w, double* __restrict__ partial,
                                               int* err) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * MB + threadIdx.x; i < ndofs; i += (int64_t)gridDim.x * MB) {
        double q[5];
        for (int v = 0; v < nd + 2; ++v) q[v] = (double)Q[v * ndofs + i];
        double m2 = q[1] * q[1];
        for (int k = 1; k < nd; ++k) m2 = m2 + q[1 + k] * q[1 + k];
        double val;
        if (kind == 0) {
            val = m2 / (2.0 * q[0]);
        } else {
            const double p = (gamma - 1.0) * (q[nd + 1] - m2 / (2.0 * q[0]));
            if (p < 0.0 || q[0] < 0.0) atomicOr(err, ERR_LOG_NEG);
            const double ent = log(p) - gamma * log(q[0]);
            val = -(q[0] * ent) / (gamma - 1.0);
        }
        const int64_t k = i / wnpe;
        s += (jac[k] * w[i - k * wnpe]) * val;
    }
    __shared__ double sh[MB / 64];
    s = mwarp_sum(s);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[wv] = s;
    __syncthreads();
    if (wv == 0) {
        s = l < MB / 64 ? sh[l] : 0.0;
        s = mwarp_sum(s);
        if (l == 0) partial[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(MB) k_monitor_finish(const double* __restrict__ partial, int n,
                                                      double* __restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += MB) s += partial[i];
    __shared__ double sh[MB / 64];
    s = mwarp_sum(s);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[wv] = s;
    __syncthreads();
    if (wv == 0) {
        s = l < MB / 64 ? sh[l] : 0.0;
        s = mwarp_sum(s);
        if (l == 0) *out = s;
    }
}

template <typename T>
hipError_t launch_zhang_shu(T* Q, int64_t ndofs, int64_t ne, int npe, int nd, int wnpe,
                            const double* jac, const double* w, const double* vol, double minval,
                            double gamma, hipStream_t s) {
    const int g = (int)((ne + LB - 1) / LB);
    if (nd == 1)
        hipLaunchKernelGGL((k_zhang_shu<T, 3>), dim3(g), dim3(LB), 0, s, Q, ndofs, ne, npe, wnpe, jac, w, vol, (T)minval, (T)gamma);
    else if (nd == 2)
        hipLaunchKernelGGL((k_zhang_shu<T, 4>), dim3(g), dim3(LB), 0, s, Q, ndofs, ne, npe, wnpe, jac, w, vol, (T)minval, (T)gamma);
    else
        hipLaunchKernelGGL((k_zhang_shu<T, 5>), dim3(g), dim3(LB), 0, s, Q, ndofs, ne, npe, wnpe, jac, w, vol, (T)minval, (T)gamma);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_monitor(const T* Q, int64_t ndofs, int wnpe, int nd, double gamma, int kind,
                          const double* jac, const double* w, double* partial, int max_blocks,
                          double* out, int* err, hipStream_t s) {
    int64_t g = (ndofs + MB - 1) / MB;
    if (g > max_blocks) g = max_blocks;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_monitor<T>, dim3((int)g), dim3(MB), 0, s, Q, ndofs, wnpe, nd, gamma, kind, jac, w,
                       partial, err);
    hipLaunchKernelGGL(k_monitor_finish, dim3(1), dim3(MB), 0, s, partial, (int)g, out);
    return hipGetLastError();
}

template hipError_t launch_zhang_shu<double>(double*, int64_t, int64_t, int, int, int, const double*,
                                             const double*, const double*, double, double, hipStream_t);
template hipError_t launch_zhang_shu<float>(float*, int64_t, int64_t, int, int, int, const double*,
                                            const double*, const double*, double, double, hipStream_t);
template hipError_t launch_monitor<double>(const double*, int64_t, int, int, double, int, const double*,
                                           const double*, double*, int, double*, int*, hipStream_t);
template hipError_t launch_monitor<float>(const float*, int64_t, int, int, double, int, const double*,
                                          const double*, double*, int, double*, int*, hipStream_t);

}  // namespace flouhip
