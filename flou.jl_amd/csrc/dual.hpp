// One-tangent forward-mode dual numbers for the device, so that the plain physics
// templates (physics_plain.hpp) differentiate themselves: the reverse product of a
// numerical flux whose derivative is not written out by hand is assembled from one
// forward evaluation per input direction.  Conventions follow the oracle's dual numbers:
// branches and comparisons take the value, |x| has derivative sign(x) with +1 at 0.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace flouhip {
namespace adj {

template <typename T>
struct Dual1 {
    T v, d;
    __device__ Dual1() : v(T(0)), d(T(0)) {}
    __device__ Dual1(double x) : v(T(x)), d(T(0)) {}
    __device__ Dual1(T x, T dx) : v(x), d(dx) {}
};

template <typename T> __device__ inline Dual1<T> operator+(Dual1<T> a, Dual1<T> b) { return {a.v + b.v, a.d + b.d}; }
template <typename T> __device__ inline Dual1<T> operator-(Dual1<T> a, Dual1<T> b) { return {a.v - b.v, a.d - b.d}; }
template <typename T> __device__ inline Dual1<T> operator-(Dual1<T> a) { return {-a.v, -a.d}; }
template <typename T> __device__ inline Dual1<T> operator*(Dual1<T> a, Dual1<T> b) {
    return {a.v * b.v, a.d * b.v + a.v * b.d};
}
template <typename T> __device__ inline Dual1<T> operator/(Dual1<T> a, Dual1<T> b) {
    const T q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
template <typename T> __device__ inline bool operator<(Dual1<T> a, Dual1<T> b) { return a.v < b.v; }
template <typename T> __device__ inline Dual1<T> sqrt(Dual1<T> a) {
    const T s = ::sqrt(a.v);
    return {s, a.d / (T(2) * s)};
}
template <typename T> __device__ inline Dual1<T> log(Dual1<T> a) { return {::log(a.v), a.d / a.v}; }
template <typename T> __device__ inline Dual1<T> fabs(Dual1<T> a) { return a.v < T(0) ? -a : a; }

}  // namespace adj
}  // namespace flouhip
