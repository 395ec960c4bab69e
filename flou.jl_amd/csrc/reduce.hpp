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

// Element-wise projection onto m^nd equispaced points (output.hip): for each of the ncols
// columns c and elements e, Qe[(c*ne + e)*m^nd + i] = ((A ⊗ .. ⊗ A) Q_e)_i with
// Q_e = Q[(c*ne + e)*n^nd + :] and A the m x n row-major interpolation matrix (device
// memory).  hipErrorInvalidValue when the generic kernel's LDS footprint exceeds 64 KB.
template <typename T>
hipError_t launch_equispaced(const T* Q, T* Qe, const double* A, int nd, int n, int m, int64_t ne, int ncols,
                             hipStream_t s);
// LDS bytes of the generic kernel and its units per block
int equispaced_lds_bytes(int nd, int n, int m, int elem_bytes, int* upb_out);

// Number of partials the reductions may write (size of the partial buffer).
int reduce_max_blocks();

}  // namespace flouhip
