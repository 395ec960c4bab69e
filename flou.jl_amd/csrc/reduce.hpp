// Launchers of the FP64 reductions (reduce.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flouhip {

template <typename T>
hipError_t launch_max_dt(const T* Q, int64_t ndofs, int npe, int nd, int eq, double gamma,
                         double anorm, double cfl, const double* dxe, double* partial,
                         double* out, int* err, hipStream_t s);

template <typename T>
hipError_t launch_integrate(const T* Q, int64_t ndofs, int npe, const double* jac,
                            const double* w, double* partial, double* out, hipStream_t s);

template <typename S, typename D>
hipError_t launch_convert(const S* src, D* dst, int64_t n, hipStream_t s);

// Zhang & Shu limiter in place (limiter.hip); the quadrature weight of node i of element e
// is jac[k]*w[g - k*wnpe] with g = e*npe + i, k = g / wnpe.
template <typename T>
hipError_t launch_zhang_shu(T* Q, int64_t ndofs, int64_t ne, int npe, int nd, int wnpe,
                            const double* jac, const double* w, const double* vol, double minval,
                            double gamma, hipStream_t s);

// Monitor Σ Jω s(Q), kind 0 kinetic energy, 1 mathematical entropy (limiter.hip).
template <typename T>
hipError_t launch_monitor(const T* Q, int64_t ndofs, int wnpe, int nd, double gamma, int kind,
                          const double* jac, const double* w, double* partial, int max_blocks,
                          double* out, int* err, hipStream_t s);

// Number of partials the reductions may write (size of the partial buffer).
int reduce_max_blocks();

}  // namespace flouhip
