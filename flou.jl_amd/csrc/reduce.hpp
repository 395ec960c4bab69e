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

// Number of partials the reductions may write (size of the partial buffer).
int reduce_max_blocks();

}  // namespace flouhip
